// Host check of the fast tile addressing (ptq2_device.h): for every lane and step of a full
// 9x9x9 tile, the element offset and the LDS offset of the fast path equal those of the
// row walk (row r = 2 it + p, y = r % 9, z = r / 9), every stored row is visited exactly once,
// and every loaded offset stays inside the tile.  Several box shapes, including the largest
// node counts for which the fast path is enabled.
#include <cstdio>
#include <cstdlib>
#include <set>

#include "ptq2_device.h"

using namespace ptq2;

static int check_box(int NX, int NY) {
  Box b{};
  b.NX = NX;
  b.NY = NY;
  b.NZ = 9;
  b.fast32 = 1;
  const long long sy = 3LL * NX, sz = sy * NY;
  const long long tile_max = 26 + 8 * sy + 8 * sz;
  std::set<long long> seen_rows;
  int bad = 0;
  for (int lane = 0; lane < BEL; lane++) {
    const FastLane L = fast_lane(b, lane);
    const int p = lane < 2 * RLEN ? lane / RLEN : 2;
    const int col = p < 2 ? lane - RLEN * p : 0;
    if (L.p != p || L.col != col) bad++;
    for (int it = 0; it < TPAIRS; it++) {
      const unsigned e = fast_eoff(b, L, it);
      if ((long long)e > tile_max) bad++;
      if (!fast_stores(L, it)) continue;
      const int r = 2 * it + p;
      const int y = r % TN, z = r / TN;
      if (z >= TN) bad++;
      int fy, fz;
      fast_row(L, it, fy, fz);
      if (fy != y || fz != z) bad++;
      const long long ref = col + sy * y + sz * z;
      if ((long long)e != ref) bad++;
      const int ii = col / 3, c = col - 3 * ii;
      const int lds = (fast_wrap(it) ? L.ldsw : L.lds0) + fast_lds_step(it);
      if (lds != c * TCS + tx(ii) + TRS * y + TPS * z) bad++;
      seen_rows.insert((long long)r * 100 + col);
    }
  }
  if ((int)seen_rows.size() != TROWS * RLEN) bad++;
  return bad;
}

int main() {
  int bad = 0;
  const int shapes[][2] = {{9, 9}, {17, 9}, {33, 65}, {257, 257}, {129, 513}, {1001, 1001}};
  for (auto &s : shapes) {
    const int e = check_box(s[0], s[1]);
    if (e) std::printf("box %d x %d: %d mismatches\n", s[0], s[1], e);
    bad += e;
  }
  if (bad) return 1;
  std::printf("tile addressing ok\n");
  return 0;
}
