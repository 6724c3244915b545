// ptq2_device.h -- device-side building blocks shared by the operator kernels.
//
// Work decomposition (MI355X, gfx950): one 64-lane wavefront owns a brick of
// 4x4x4 Q2 elements, lane = element.  The brick's 9x9x9 node tile of
// coordinates and velocity is staged once in LDS with row-contiguous global
// loads; every lane then runs the whole sum-factorized element computation in
// registers (no cross-lane traffic), accumulates its 81 outputs into an LDS
// output tile with ds_add_f64 (each instruction's 64 addresses are distinct,
// so the sum order is the fixed program order), and the tile is written back
// with row-contiguous stores.  All bricks run in one launch: a brick stores the
// nodes it owns and parks its shared upper-face values in a face slot that a
// second pass adds into the owners (store_tile_own).  The older path, kept for
// callers without a face buffer, launches bricks in 8 colours (brick parity in
// x, y, z) so that bricks of one launch share no node (store_tile).  Either way
// the sum order is fixed and the result is bitwise reproducible.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptq2 {

constexpr int BE = 4;               // elements per brick edge
constexpr int BEL = BE * BE * BE;   // elements per brick (= lanes per wave)
constexpr int TN = 2 * BE + 1;      // tile nodes per edge
constexpr int TROWS = TN * TN;      // rows per tile
constexpr int RLEN = 3 * TN;        // values per global row (interlaced nodes)

// LDS tile layout, one array per component: value (x,y,z) at
//   c*TCS + tx(x) + TRS*y + TPS*z,   tx(x) = x/2 for even x, 5 + x/2 for odd x.
// Element (lx,ly,lz) of a lane reads its node (i,j,k) at lane offset
//   lx + 2*TRS*ly + 2*TPS*lz  plus  xo(i) + TRS*j + TPS*k,  xo = {0,5,1}.
// With TRS = 10, TPS = 92 and the lane -> lz order 0,2,1,3 (lane_elem), the 64
// lane offsets of every 8-byte read fall in distinct banks within each group of 16
// lanes (32 banks) and of 32 lanes (64 banks): the element loops are free of LDS
// bank conflicts.  Two FP64 tiles take 39744 B, so four bricks fit a CU.
constexpr int TRS = 10;             // row stride
constexpr int TPS = 92;             // plane stride
constexpr int TCS = TN * TPS;       // component stride
__host__ __device__ constexpr int tx(int x) { return (x & 1) ? 5 + (x >> 1) : (x >> 1); }
__host__ __device__ constexpr int xo(int i) { return i == 0 ? 0 : (i == 1 ? 5 : 1); }

// Mesh box: NX x NY x NZ nodes, mx x my x mz owned elements, first node of
// element (0,0,0) at (i0,j0,k0); bricks nbx x nby x nbz.
struct Box {
  int mx, my, mz;
  int NX, NY, NZ;
  int i0, j0, k0;
  int nbx, nby, nbz;
  int fast32;  // element offsets inside a 9^3-node tile fit 32 bits (the fast tile paths)
};

// 1D Q2 Lagrange basis on nodes {-1,0,1} at the Gauss points {-sqrt(3/5),0,sqrt(3/5)}:
// B(q,n) = l_n(x_q), D(q,n) = l_n'(x_q) (src/stokes_operators_tensor.c:193-206).
__host__ __device__ constexpr double gx(int q) { return q == 0 ? -0.77459666924148337704 : (q == 1 ? 0.0 : 0.77459666924148337704); }
__host__ __device__ constexpr double gB(int q, int n) {
  return n == 0 ? 0.5 * (gx(q) * gx(q) - gx(q)) : (n == 1 ? 1.0 - gx(q) * gx(q) : 0.5 * (gx(q) * gx(q) + gx(q)));
}
__host__ __device__ constexpr double gD(int q, int n) {
  return n == 0 ? gx(q) - 0.5 : (n == 1 ? -2.0 * gx(q) : gx(q) + 0.5);
}
__host__ __device__ constexpr double gw1(int q) { return q == 1 ? 8.0 / 9.0 : 5.0 / 9.0; }
// The exact zeros of the tables: B(1,0)=B(1,2)=0 and D(1,1)=0.
__host__ __device__ constexpr bool nzB(int q, int n) { return !(q == 1 && n != 1); }
__host__ __device__ constexpr bool nzD(int q, int n) { return !(q == 1 && n == 1); }
__host__ __device__ constexpr bool isoneB(int q, int n) { return q == 1 && n == 1; }

template <typename T>
__device__ __forceinline__ T fmac(double c, T x, T acc) {
  return fma(static_cast<T>(c), x, acc);
}

// brick of a coloured launch
struct BrickId {
  int bx, by, bz, lin;
};
__device__ __forceinline__ BrickId brick_of(const Box &b, int color, int bid) {
  int cx = color & 1, cy = (color >> 1) & 1, cz = (color >> 2) & 1;
  int ncx = (b.nbx - cx + 1) >> 1, ncy = (b.nby - cy + 1) >> 1;
  BrickId r;
  r.bx = cx + 2 * (bid % ncx);
  r.by = cy + 2 * ((bid / ncx) % ncy);
  r.bz = cz + 2 * (bid / (ncx * ncy));
  r.lin = r.bx + b.nbx * (r.by + b.nby * r.bz);
  return r;
}
// brick of an uncoloured launch (outputs not shared between bricks)
__device__ __forceinline__ BrickId brick_lin(const Box &b, int bid) {
  BrickId r;
  r.bx = bid % b.nbx;
  r.by = (bid / b.nbx) % b.nby;
  r.bz = bid / (b.nbx * b.nby);
  r.lin = bid;
  return r;
}

// brick of an uncoloured launch, XCD-aware: workgroups are dealt to the 8 XCDs in turn
// (workgroup w runs on XCD w % 8), and XCD k takes the k-th contiguous eighth of the bricks
// in x-fastest order, so bricks that share tile faces mostly run on one L2
__device__ __forceinline__ BrickId brick_xcd(const Box &b, int bid) {
  const int nb = b.nbx * b.nby * b.nbz;
  const int q = nb >> 3, r = nb & 7, k = bid & 7, i = bid >> 3;
  return brick_lin(b, k * q + min(k, r) + i);
}

struct Tile {
  int ex0, ey0, ez0;   // first element of the brick
  int nex, ney, nez;   // elements in the brick per direction
  int64_t node0;       // box node index of tile node (0,0,0)
  int tnx, tny, tnz;   // tile nodes present per direction
};
__device__ __forceinline__ Tile tile_of(const Box &b, const BrickId &br) {
  Tile t;
  t.ex0 = br.bx * BE; t.ey0 = br.by * BE; t.ez0 = br.bz * BE;
  t.nex = min(BE, b.mx - t.ex0); t.ney = min(BE, b.my - t.ey0); t.nez = min(BE, b.mz - t.ez0);
  t.tnx = 2 * t.nex + 1; t.tny = 2 * t.ney + 1; t.tnz = 2 * t.nez + 1;
  t.node0 = (int64_t)(b.i0 + 2 * t.ex0) + (int64_t)b.NX * ((int64_t)(b.j0 + 2 * t.ey0) + (int64_t)b.NY * (int64_t)(b.k0 + 2 * t.ez0));
  return t;
}

// Tile traffic.  A tile is 81 rows (y,z) of 27 values (3 components of 9 nodes), and
// a row is a contiguous run of 27 values in global memory.  Lanes 0-53 walk the rows
// in pairs: lane = 27 p + col takes column col of row 2 it + p at step it (41 steps).
// Each lane's column terms are computed once; a step costs a few integer adds.  All
// loads of a pass are issued before the first LDS store (one exposed memory latency
// per tile instead of one per value).
constexpr int TPAIRS = (TROWS + 1) / 2;
struct TileLane {
  int p;           // row parity of this lane (0/1), 2 = idle lane
  int lds_col;     // LDS offset of the column: c*TCS + tx(ii)
  bool col_ok;     // column inside the box
  bool shx;        // node column shared with the x-neighbour brick
  int64_t gcol;    // global index of the column in row (0,0)
  int64_t sy, sz;  // global strides of a tile row in y and z
};
__device__ __forceinline__ TileLane tile_lane(const Box &b, const Tile &t, const BrickId &br, int lane) {
  TileLane L;
  L.p = lane < 2 * RLEN ? lane / RLEN : 2;
  const int col = lane - RLEN * (lane < 2 * RLEN ? L.p : 0);
  const int ii = col / 3, c = col - 3 * ii;
  L.lds_col = c * TCS + tx(ii);
  L.col_ok = L.p < 2 && ii < t.tnx;
  L.shx = (ii == 0 && br.bx > 0) || (ii == 2 * BE && br.bx < b.nbx - 1);
  // the global column is clamped into the box, so every load address is valid
  const int colc = L.p < 2 ? 3 * min(ii, t.tnx - 1) + c : 0;
  L.gcol = 3 * t.node0 + colc;
  L.sy = 3 * (int64_t)b.NX;
  L.sz = 3 * (int64_t)b.NX * (int64_t)b.NY;
  return L;
}
// row (y,z) of step it for this lane
__device__ __forceinline__ void tile_row(int it, const TileLane &L, int &y, int &z) {
  y = (2 * it) % TN + L.p;
  z = (2 * it) / TN;
  if (y >= TN) {
    y -= TN;
    z += 1;
  }
}
// global index of row (y,z) clamped into the box (always a valid address)
__device__ __forceinline__ int64_t tile_gidx(const Tile &t, const TileLane &L, int y, int z) {
  return L.gcol + L.sy * min(y, t.tny - 1) + L.sz * min(z, t.tnz - 1);
}

// Fast tile traffic for a full 9x9x9 tile (every brick of a mesh whose element counts are
// multiples of 4, and the interior bricks of any mesh).  The tile's first node gives one
// wave-uniform base pointer per field; a lane addresses its rows with 32-bit element offsets
// from it.  Row r = 2 it + p of the tile (y = r % 9, z = r / 9) sits at element offset
//   sy y + sz z = sy r + (sz - 9 sy) (r / 9)          (sy = 3 NX, sz = 3 NX NY)
// and its column at LDS offset TRS y + TPS z = 10 r + 2 (r / 9) (TPS = 9 TRS + 2).  For a
// lane of parity p, (2 it + p) / 9 = (2 it) / 9 except on the odd lanes of the steps with
// (2 it) % 9 == 8, whose row wraps to the next plane: those steps use a second per-lane base.
// All per-step terms are then compile-time or wave-uniform, so a step costs one 32-bit add
// per access instead of a 64-bit multiply chain, and the LDS offsets are instruction
// immediates.  The last step's odd lanes (row 81) and the idle lanes 54-63 read row 0 and
// store nothing.
__host__ __device__ __forceinline__ bool tile_fast(const Box &b, const Tile &t) {
  return b.fast32 && t.tnx == TN && t.tny == TN && t.tnz == TN;
}
struct FastLane {
  int p;              // row parity 0/1, 2 = idle lane
  int col;            // column inside the row (0..26)
  int ii, c;          // node column and component
  int lds0, ldsw;     // LDS offset of row (p,0); the wrap-step base
  uint32_t e0, ew;    // element offset of row (p,0); the wrap-step base
};
__host__ __device__ __forceinline__ FastLane fast_lane(const Box &b, int lane) {
  FastLane L;
  L.p = lane < 2 * RLEN ? lane / RLEN : 2;
  L.col = L.p < 2 ? lane - RLEN * L.p : 0;
  L.ii = L.col / 3;
  L.c = L.col - 3 * L.ii;
  const uint32_t sy = 3u * (uint32_t)b.NX, sz = sy * (uint32_t)b.NY;
  const int odd = L.p == 1;
  L.lds0 = L.c * TCS + tx(L.ii) + (odd ? TRS : 0);
  L.ldsw = L.lds0 + (odd ? 2 : 0);
  L.e0 = (uint32_t)L.col + (odd ? sy : 0u);
  L.ew = L.e0 + (odd ? sz - 9u * sy : 0u);
  return L;
}
// uniform element offset of step it for the even row 2 it
__host__ __device__ __forceinline__ uint32_t fast_step(const Box &b, int it) {
  const uint32_t sy = 3u * (uint32_t)b.NX, sz = sy * (uint32_t)b.NY;
  return 2u * (uint32_t)it * sy + (sz - 9u * sy) * (uint32_t)((2 * it) / 9);
}
__host__ __device__ __forceinline__ constexpr int fast_lds_step(int it) { return 20 * it + 2 * ((2 * it) / 9); }
__host__ __device__ __forceinline__ constexpr bool fast_wrap(int it) { return (2 * it) % 9 == 8; }
// this lane's element offset at step it (row 0 for the lanes that store nothing at that step)
__host__ __device__ __forceinline__ uint32_t fast_eoff(const Box &b, const FastLane &L, int it) {
  if (it == TPAIRS - 1 && L.p == 1) return (uint32_t)L.col;
  return (fast_wrap(it) ? L.ew : L.e0) + fast_step(b, it);
}
// (y, z) of this lane's row at step it
__host__ __device__ __forceinline__ void fast_row(const FastLane &L, int it, int &y, int &z) {
  const int odd = L.p == 1, y0 = (2 * it) % TN, z0 = (2 * it) / TN;
  y = fast_wrap(it) ? (odd ? 0 : y0) : y0 + odd;
  z = fast_wrap(it) ? z0 + odd : z0;
}
__host__ __device__ __forceinline__ bool fast_stores(const FastLane &L, int it) {
  return L.p == 0 || (L.p == 1 && it < TPAIRS - 1);
}
// element e of a field from a wave-uniform base: the byte offset is formed in 32 bits, so the
// access is one global instruction with a scalar base and a 32-bit vector offset
template <typename S>
__device__ __forceinline__ S &fast_at(S *base, uint32_t e) {
  const uint32_t bo = e * (uint32_t)sizeof(S);
  return *reinterpret_cast<S *>(reinterpret_cast<char *>(base) + bo);
}
template <typename S>
__device__ __forceinline__ const S &fast_at(const S *base, uint32_t e) {
  const uint32_t bo = e * (uint32_t)sizeof(S);
  return *reinterpret_cast<const S *>(reinterpret_cast<const char *>(base) + bo);
}

// two fields, the second masked (Dirichlet dofs read as zero) if HM
template <bool HM, typename T, typename S>
__device__ __forceinline__ void load_tile2_fast(const Box &b, const Tile &t, const S *__restrict__ g1,
                                                T *__restrict__ s1, const S *__restrict__ g2,
                                                const uint8_t *__restrict__ mask2, T *__restrict__ s2, int lane) {
  const FastLane L = fast_lane(b, lane);
  const S *b1 = g1 + 3 * t.node0, *b2 = g2 + 3 * t.node0;
  const uint8_t *bm = HM ? mask2 + 3 * t.node0 : nullptr;
  S v1[TPAIRS], v2[TPAIRS];
  int mk[TPAIRS];
#pragma unroll
  for (int it = 0; it < TPAIRS; it++) {
    const uint32_t e = fast_eoff(b, L, it);
    v1[it] = fast_at(b1, e);
    v2[it] = fast_at(b2, e);
    mk[it] = HM ? (int)fast_at(bm, e) : 0;
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int it = 0; it < TPAIRS; it++) {
    if (fast_stores(L, it)) {
      const int o = (fast_wrap(it) ? L.ldsw : L.lds0) + fast_lds_step(it);
      s1[o] = static_cast<T>(v1[it]);
      s2[o] = mk[it] ? T(0) : static_cast<T>(v2[it]);
    }
  }
}
// one field, masked if HM
template <bool HM, typename T, typename S>
__device__ __forceinline__ void load_tile_fast(const Box &b, const Tile &t, const S *__restrict__ g,
                                               const uint8_t *__restrict__ mask, T *__restrict__ s, int lane) {
  const FastLane L = fast_lane(b, lane);
  const S *b1 = g + 3 * t.node0;
  const uint8_t *bm = HM ? mask + 3 * t.node0 : nullptr;
  S v[TPAIRS];
  int mk[TPAIRS];
#pragma unroll
  for (int it = 0; it < TPAIRS; it++) {
    const uint32_t e = fast_eoff(b, L, it);
    v[it] = fast_at(b1, e);
    mk[it] = HM ? (int)fast_at(bm, e) : 0;
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int it = 0; it < TPAIRS; it++) {
    if (fast_stores(L, it)) {
      const int o = (fast_wrap(it) ? L.ldsw : L.lds0) + fast_lds_step(it);
      s[o] = mk[it] ? T(0) : static_cast<T>(v[it]);
    }
  }
}

// Stage an interlaced 3-component nodal field into an LDS tile.  When mask is given,
// Dirichlet dofs read as zero (BCListInsertLocalZero, src/dmda_bcs.c:581-608).  Tile
// nodes outside the box read as zero.
template <bool HM, typename T, typename S>
__device__ __forceinline__ void load_tile_m(const Box &b, const Tile &t, const BrickId &br, const S *__restrict__ g,
                                            const uint8_t *__restrict__ mask, T *__restrict__ s, int lane) {
  const TileLane L = tile_lane(b, t, br, lane);
  // all loads first (addresses clamped into the box, so none is conditional), then
  // the selects and LDS stores
  S v[TPAIRS];
  int mk[TPAIRS];  // full registers: packed byte loads would serialize
#pragma unroll
  for (int it = 0; it < TPAIRS; it++) {
    int y, z;
    tile_row(it, L, y, z);
    const int64_t gi = tile_gidx(t, L, y, z);
    v[it] = g[gi];
    mk[it] = HM ? (int)mask[gi] : 0;
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int it = 0; it < TPAIRS; it++) {
    int y, z;
    tile_row(it, L, y, z);
    const bool ok = L.col_ok && y < t.tny && z < t.tnz;
    if (L.p < 2 && z < TN) s[L.lds_col + TRS * y + TPS * z] = (ok && !mk[it]) ? static_cast<T>(v[it]) : T(0);
  }
}

template <typename T, typename S>
__device__ __forceinline__ void load_tile(const Box &b, const Tile &t, const BrickId &br, const S *__restrict__ g,
                                          const uint8_t *__restrict__ mask, T *__restrict__ s, int lane) {
  if (tile_fast(b, t)) {
    if (mask)
      load_tile_fast<true>(b, t, g, mask, s, lane);
    else
      load_tile_fast<false>(b, t, g, mask, s, lane);
    return;
  }
  if (mask)
    load_tile_m<true>(b, t, br, g, mask, s, lane);
  else
    load_tile_m<false>(b, t, br, g, mask, s, lane);
}

// Two fields over the same tile in one pass; mask2 applies to the second (if HM).
template <bool HM, typename T, typename S>
__device__ __forceinline__ void load_tile2_m(const Box &b, const Tile &t, const BrickId &br, const S *__restrict__ g1,
                                             T *__restrict__ s1, const S *__restrict__ g2,
                                             const uint8_t *__restrict__ mask2, T *__restrict__ s2, int lane) {
  const TileLane L = tile_lane(b, t, br, lane);
  S v1[TPAIRS], v2[TPAIRS];
  int mk[TPAIRS];
#pragma unroll
  for (int it = 0; it < TPAIRS; it++) {
    int y, z;
    tile_row(it, L, y, z);
    const int64_t gi = tile_gidx(t, L, y, z);
    v1[it] = g1[gi];
    v2[it] = g2[gi];
    mk[it] = HM ? (int)mask2[gi] : 0;
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int it = 0; it < TPAIRS; it++) {
    int y, z;
    tile_row(it, L, y, z);
    const bool ok = L.col_ok && y < t.tny && z < t.tnz;
    if (L.p < 2 && z < TN) {
      const int o = L.lds_col + TRS * y + TPS * z;
      s1[o] = ok ? static_cast<T>(v1[it]) : T(0);
      s2[o] = (ok && !mk[it]) ? static_cast<T>(v2[it]) : T(0);
    }
  }
}

template <typename T, typename S>
__device__ __forceinline__ void load_tile2(const Box &b, const Tile &t, const BrickId &br, const S *__restrict__ g1,
                                           T *__restrict__ s1, const S *__restrict__ g2,
                                           const uint8_t *__restrict__ mask2, T *__restrict__ s2, int lane) {
  if (tile_fast(b, t)) {
    if (mask2)
      load_tile2_fast<true>(b, t, g1, s1, g2, mask2, s2, lane);
    else
      load_tile2_fast<false>(b, t, g1, s1, g2, mask2, s2, lane);
    return;
  }
  if (mask2)
    load_tile2_m<true>(b, t, br, g1, s1, g2, mask2, s2, lane);
  else
    load_tile2_m<false>(b, t, br, g1, s1, g2, mask2, s2, lane);
}

// Write an LDS output tile back.  Bricks of one colour never share a node.  In
// overwrite mode the first brick to touch a node (parity 0 along every axis on
// which the node is shared, i.e. the lowest colour) stores, the later colours add.
// The global values to add to are all loaded before the first store.  The LDS tile may
// be wider than the vector (an FP64 tile accumulating an FP32 operator): each node is
// summed in the tile's type and rounded once.
template <typename S, typename T>
__device__ __forceinline__ void store_tile(const Box &b, const BrickId &br, const Tile &t, const S *__restrict__ s,
                                           T *__restrict__ g, bool add_mode, int lane) {
  // opaque copy of the lane: keeps the compiler from sharing this tile's address
  // arithmetic with the loads at the top of the kernel (which would keep it live
  // across the whole element computation)
  asm volatile("" : "+v"(lane));
  const TileLane L = tile_lane(b, t, br, lane);
  T v[TPAIRS];
#pragma unroll
  for (int it = 0; it < TPAIRS; it++) {
    int y, z;
    tile_row(it, L, y, z);
    const bool ok = L.col_ok && y < t.tny && z < t.tnz;
    const bool shy = (y == 0 && br.by > 0) || (y == 2 * BE && br.by < b.nby - 1);
    const bool shz = (z == 0 && br.bz > 0) || (z == 2 * BE && br.bz < b.nbz - 1);
    const bool first = (!L.shx || !(br.bx & 1)) && (!shy || !(br.by & 1)) && (!shz || !(br.bz & 1));
    const bool rd = ok && (add_mode || !first);
    const T val = g[tile_gidx(t, L, y, z)];
    v[it] = rd ? val : T(0);
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int it = 0; it < TPAIRS; it++) {
    int y, z;
    tile_row(it, L, y, z);
    if (L.col_ok && y < t.tny && z < t.tnz) g[tile_gidx(t, L, y, z)] = T(v[it] + s[L.lds_col + TRS * y + TPS * z]);
  }
}

// Uncoloured output, owner writes: every tile node has one owning brick, the brick whose
// local coordinates of the node are all below 2*BE (or, for the last brick of an axis, that
// axis' upper boundary node).  The brick stores (or, in add mode, adds) the nodes it owns
// straight into g, and the nodes of its upper faces that a neighbour owns into its face slot
// of FACEVALS values: three TN x TN planes of interlaced 3-vectors, X [z][y] (x = 2*BE),
// Y [z][x] (y = 2*BE), Z [y][x] (z = 2*BE), a node going to the first of X, Y, Z that
// applies.  k_gather_faces adds the slots into their owners.
constexpr int TILEVALS = 3 * TN * TN * TN;
constexpr int FACEVALS = 3 * 3 * TN * TN;
// One Chebyshev step fused into the apply's store (ChebychevRN, src/ksp_chebyrn.c:344-411):
// out = cr dinv (b - y) + ck pkm1 + cx x (pkm1 null: cr dinv (b - y) + x) with y = (A x)_i,
// y = x on Dirichlet dofs (MatMult semantics).  The explicit FMA order is shared by the
// apply's store and the face pass, so a dof gets the same bits whichever of them updates it.
template <typename T>
struct ChebOut {
  const T *b, *dinv, *x, *pkm1;
  T *out;
  const uint8_t *mask;
  T cr, ck, cx;
};
// The update from its operands.  Contraction is off: the compiler would otherwise fuse v + x into
// an FMA at some call sites and not at others, and the two sites must agree bit for bit.
template <typename T>
__host__ __device__ __forceinline__ T cheb_value(const ChebOut<T> &co, T d, T bv, T y, T x, T pk, bool has_p) {
#pragma clang fp contract(off)
  const T v = (co.cr * d) * (bv - y);
  return has_p ? fma(co.cx, x, fma(co.ck, pk, v)) : v + x;
}
template <typename T>
__host__ __device__ __forceinline__ T cheb_update(const ChebOut<T> &co, int64_t i, T y) {
  return cheb_value(co, co.dinv[i], co.b[i], y, co.x[i], co.pkm1 ? co.pkm1[i] : T(0), co.pkm1 != nullptr);
}

// the full-tile case with the addressing of load_tile2_fast
template <bool CH, typename S, typename T>
__device__ __forceinline__ void store_tile_own_fast(const Box &b, const BrickId &br, const Tile &t,
                                                    const S *__restrict__ s, T *__restrict__ g,
                                                    T *__restrict__ face, bool add_mode, int lane,
                                                    const ChebOut<T> *co) {
  const FastLane L = fast_lane(b, lane);
  const bool ownx = L.ii < 2 * BE || br.bx == b.nbx - 1;
  const bool lasty = br.by == b.nby - 1, lastz = br.bz == b.nbz - 1;
  T *gb = g + 3 * t.node0;
  T *fo = face + (int64_t)br.lin * FACEVALS + L.c;
  if constexpr (CH) {
    // the update's operands of a chunk of steps are all loaded before the first is used (one
    // exposed memory latency per chunk, not per step); every tile address is inside the box
    constexpr int CHUNK = sizeof(T) == 4 ? TPAIRS : 14;
    const T *bb = co->b + 3 * t.node0, *bd = co->dinv + 3 * t.node0, *bx = co->x + 3 * t.node0;
    const T *bp = co->pkm1 ? co->pkm1 + 3 * t.node0 : nullptr;
    const uint8_t *bm = co->mask ? co->mask + 3 * t.node0 : nullptr;
    T *bo = co->out + 3 * t.node0;
#pragma unroll
    for (int c0 = 0; c0 < TPAIRS; c0 += CHUNK) {
      T vb[CHUNK], vd[CHUNK], vx[CHUNK], vp[CHUNK];
      int vm[CHUNK];
#pragma unroll
      for (int k = 0; k < CHUNK; k++) {
        const int it = c0 + k;
        if (it >= TPAIRS) break;
        const uint32_t e = fast_eoff(b, L, it);
        vb[k] = fast_at(bb, e);
        vd[k] = fast_at(bd, e);
        vx[k] = fast_at(bx, e);
        vp[k] = bp ? fast_at(bp, e) : T(0);
        vm[k] = bm ? (int)fast_at(bm, e) : 0;
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < CHUNK; k++) {
        const int it = c0 + k;
        if (it >= TPAIRS) break;
        if (!fast_stores(L, it)) continue;
        int y, z;
        fast_row(L, it, y, z);
        const bool owny = y < 2 * BE || lasty, ownz = z < 2 * BE || lastz;
        const S val = s[(fast_wrap(it) ? L.ldsw : L.lds0) + fast_lds_step(it)];
        const uint32_t e = fast_eoff(b, L, it);
        if (ownx && owny && ownz) {
          const bool lower = (L.ii == 0 && br.bx > 0) || (y == 0 && br.by > 0) || (z == 0 && br.bz > 0);
          if (vm[k] || !lower) {
            fast_at(bo, e) = cheb_value(*co, vd[k], vb[k], vm[k] ? vx[k] : T(val), vx[k], vp[k], bp != nullptr);
          } else {
            fast_at(gb, e) = T(val);
          }
        } else {
          const int slot = !ownx ? z * TN + y : (!owny ? TN * TN + z * TN + L.ii : 2 * TN * TN + y * TN + L.ii);
          fo[3 * slot] = T(val);
        }
      }
    }
    return;
  }
  T v[TPAIRS];
  if (add_mode) {
#pragma unroll
    for (int it = 0; it < TPAIRS; it++) v[it] = fast_at(gb, fast_eoff(b, L, it));
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int it = 0; it < TPAIRS; it++) {
    if (!fast_stores(L, it)) continue;
    int y, z;
    fast_row(L, it, y, z);
    const bool owny = y < 2 * BE || lasty, ownz = z < 2 * BE || lastz;
    const S val = s[(fast_wrap(it) ? L.ldsw : L.lds0) + fast_lds_step(it)];
    if (ownx && owny && ownz) {
      fast_at(gb, fast_eoff(b, L, it)) = add_mode ? T(v[it] + val) : T(val);
    } else {
      const int slot = !ownx ? z * TN + y : (!owny ? TN * TN + z * TN + L.ii : 2 * TN * TN + y * TN + L.ii);
      fo[3 * slot] = T(val);
    }
  }
}

// CH: the Chebyshev update of co is applied to every owned node that is final here (see
// store_tile_own_fast); the others are stored into g for the face pass.
template <bool CH = false, typename S, typename T>
__device__ __forceinline__ void store_tile_own(const Box &b, const BrickId &br, const Tile &t, const S *__restrict__ s,
                                               T *__restrict__ g, T *__restrict__ face, bool add_mode, int lane,
                                               const ChebOut<T> *co = nullptr) {
  asm volatile("" : "+v"(lane));  // see store_tile
  if (tile_fast(b, t)) {
    store_tile_own_fast<CH>(b, br, t, s, g, face, add_mode, lane, co);
    return;
  }
  const TileLane L = tile_lane(b, t, br, lane);
  const int col = lane - RLEN * (L.p < 2 ? L.p : 0);
  const int ii = col / 3, c = col - 3 * ii;
  const bool ownx = ii < 2 * BE || br.bx == b.nbx - 1;
  const bool lasty = br.by == b.nby - 1, lastz = br.bz == b.nbz - 1;
  T *fo = face + (int64_t)br.lin * FACEVALS + c;
  T v[TPAIRS];
  if (add_mode) {
#pragma unroll
    for (int it = 0; it < TPAIRS; it++) {
      int y, z;
      tile_row(it, L, y, z);
      v[it] = g[tile_gidx(t, L, y, z)];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int it = 0; it < TPAIRS; it++) {
    int y, z;
    tile_row(it, L, y, z);
    if (!(L.col_ok && y < t.tny && z < t.tnz)) continue;
    const bool owny = y < 2 * BE || lasty, ownz = z < 2 * BE || lastz;
    const S val = s[L.lds_col + TRS * y + TPS * z];
    if (ownx && owny && ownz) {
      const int64_t gi = tile_gidx(t, L, y, z);
      if constexpr (CH) {
        const bool dir = co->mask && co->mask[gi];
        const bool lower = (ii == 0 && br.bx > 0) || (y == 0 && br.by > 0) || (z == 0 && br.bz > 0);
        if (dir || !lower)
          co->out[gi] = cheb_update(*co, gi, dir ? co->x[gi] : T(val));
        else
          g[gi] = T(val);
      } else {
        g[gi] = add_mode ? T(v[it] + val) : T(val);
      }
    } else {
      const int slot = !ownx ? z * TN + y : (!owny ? TN * TN + z * TN + ii : 2 * TN * TN + y * TN + ii);
      fo[3 * slot] = T(val);
    }
  }
}

template <typename T>
__device__ __forceinline__ void zero_tile(T *__restrict__ s, int lane) {
  for (int f = lane; f < 3 * TCS; f += BEL) s[f] = T(0);
}

__device__ __forceinline__ void lds_add(double *p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add(float *p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// lane -> element (lx,ly,lz) inside the brick; lanes 16-31 take lz = 2 and
// lanes 32-47 lz = 1 (bank-conflict-free tile reads, see the layout above).
struct LaneElem {
  int lx, ly, lz;
};
__host__ __device__ __forceinline__ LaneElem lane_elem(int lane) {
  const int g = lane >> 4;
  return LaneElem{lane & (BE - 1), (lane >> 2) & (BE - 1), ((g & 1) << 1) | (g >> 1)};
}
// inverse of lane_elem: the lane of element (lx,ly,lz) of a brick (the z order swap is
// its own inverse)
__host__ __device__ __forceinline__ int elem_lane(int lx, int ly, int lz) {
  return lx + BE * ly + BE * BE * (((lz & 1) << 1) | (lz >> 1));
}
__device__ __forceinline__ int lane_base(int lane) {
  const LaneElem e = lane_elem(lane);
  return e.lx + 2 * TRS * e.ly + 2 * TPS * e.lz;
}
__device__ __forceinline__ bool lane_active(const Tile &t, int lane) {
  const LaneElem e = lane_elem(lane);
  return e.lx < t.nex && e.ly < t.ney && e.lz < t.nez;
}
__device__ __forceinline__ int64_t lane_element(const Box &b, const Tile &t, int lane) {
  const LaneElem e = lane_elem(lane);
  return (int64_t)(t.ex0 + e.lx) + (int64_t)b.mx * ((int64_t)(t.ey0 + e.ly) + (int64_t)b.my * (int64_t)(t.ez0 + e.lz));
}

// offset of element-local node (i,j,k), component c, from the lane base
__host__ __device__ constexpr int toff(int c, int i, int j, int k) { return c * TCS + xo(i) + j * TRS + k * TPS; }

// Tile value at element-local node (i,j,k), component c.
template <typename T>
__device__ __forceinline__ T tv(const T *s, int base, int c, int i, int j, int k) {
  return s[base + toff(c, i, j, k)];
}

// Row data of one field for quadrature row (qb,qc): for each component c and
// element-local node column i,
//   r0[c][i] = sum_jk B(qb,j) B(qc,k) F,  r1 = sum D(qb,j) B(qc,k) F,  r2 = sum B(qb,j) D(qc,k) F.
template <int qb, int qc, typename T>
__device__ __forceinline__ void row_data(const T *s, int base, T r0[3][3], T r1[3][3], T r2[3][3]) {
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      T a0 = T(0), a1 = T(0), a2 = T(0);
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          const bool u0 = nzB(qb, j) && nzB(qc, k);
          const bool u1 = nzD(qb, j) && nzB(qc, k);
          const bool u2 = nzB(qb, j) && nzD(qc, k);
          if (u0 || u1 || u2) {
            T v = tv(s, base, c, i, j, k);
            if (u0) a0 = fmac(gB(qb, j) * gB(qc, k), v, a0);
            if (u1) a1 = fmac(gD(qb, j) * gB(qc, k), v, a1);
            if (u2) a2 = fmac(gB(qb, j) * gD(qc, k), v, a2);
          }
        }
      r0[c][i] = a0; r1[c][i] = a1; r2[c][i] = a2;
    }
}

// Reference-gradient at quadrature point (qa,qb,qc) from the row data:
// g[d][c] = d F_c / d xi_d.
template <int qa, typename T>
__device__ __forceinline__ void qp_grad(const T r0[3][3], const T r1[3][3], const T r2[3][3], T g[3][3]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    T a0 = T(0), a1 = T(0), a2 = T(0);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      if (nzD(qa, i)) a0 = fmac(gD(qa, i), r0[c][i], a0);
      if (nzB(qa, i)) a1 = fmac(gB(qa, i), r1[c][i], a1);
      if (nzB(qa, i)) a2 = fmac(gB(qa, i), r2[c][i], a2);
    }
    g[0][c] = a0; g[1][c] = a1; g[2][c] = a2;
  }
}

// value at quadrature point (qa,...) from the B-B row data r0
template <int qa, typename T>
__device__ __forceinline__ void qp_value(const T r0[3][3], T v[3]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    T a = T(0);
#pragma unroll
    for (int i = 0; i < 3; i++)
      if (nzB(qa, i)) a = fmac(gB(qa, i), r0[c][i], a);
    v[c] = a;
  }
}

// Adjugate of J (J[d][l] = dx_l/dxi_d) arranged so that inv(J)[k][m] = A[k][m]/det,
// i.e. A[k][m] = d xi_m / d x_k * det.  Same cofactors as JacobianInvertNEV
// (src/stokes_operators_tensor.c:94-125).
template <typename T>
__device__ __forceinline__ T adjugate(const T J[3][3], T A[3][3]) {
  A[0][0] = J[1][1] * J[2][2] - J[2][1] * J[1][2];
  A[1][0] = -(J[1][0] * J[2][2] - J[2][0] * J[1][2]);
  A[2][0] = J[1][0] * J[2][1] - J[2][0] * J[1][1];
  A[0][1] = -(J[0][1] * J[2][2] - J[2][1] * J[0][2]);
  A[1][1] = J[0][0] * J[2][2] - J[2][0] * J[0][2];
  A[2][1] = -(J[0][0] * J[2][1] - J[2][0] * J[0][1]);
  A[0][2] = J[0][1] * J[1][2] - J[1][1] * J[0][2];
  A[1][2] = -(J[0][0] * J[1][2] - J[1][0] * J[0][2]);
  A[2][2] = J[0][0] * J[1][1] - J[1][0] * J[0][1];
  return J[0][0] * A[0][0] + J[0][1] * A[1][0] + J[0][2] * A[2][0];
}

// FP64 reciprocal: hardware estimate refined by two Newton steps (error ~1 ulp).
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ float rcp32(float x) { return 1.0f / x; }

// Stress action at one quadrature point.  With M[k][l] = sum_m g[m][l] A[k][m]
// (= det * du_l/dx_k) and S = s (M + M^T), s = w eta / det, the test-function
// coefficient is dv[k][l] = sum_m S[m][l] A[m][k]; this equals w det J^-T (2 eta eps(u))
// of QuadratureAction (src/stokes_operators_tensor.c:127-179).
template <typename T>
__device__ __forceinline__ void stress_action(const T A[3][3], const T g[3][3], T s, T dv[3][3]) {
  T M[3][3];
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int l = 0; l < 3; l++) M[k][l] = g[0][l] * A[k][0] + g[1][l] * A[k][1] + g[2][l] * A[k][2];
  T S[3][3];
  T s2 = s + s;
  S[0][0] = s2 * M[0][0];
  S[1][1] = s2 * M[1][1];
  S[2][2] = s2 * M[2][2];
  S[0][1] = S[1][0] = s * (M[0][1] + M[1][0]);
  S[0][2] = S[2][0] = s * (M[0][2] + M[2][0]);
  S[1][2] = S[2][1] = s * (M[1][2] + M[2][1]);
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int l = 0; l < 3; l++) dv[k][l] = S[0][l] * A[0][k] + S[1][l] * A[1][k] + S[2][l] * A[2][k];
}

// The same with the pressure of the full block: S -= wp I (wp = w p_q), which adds
// -w p_q adj(J)^T, the A12 term of QuadratureAction_A_AVX (src/stokes_operators_avx.c:376-437);
// returns trace(M) = det div u for the A21 term.
template <typename T>
__device__ __forceinline__ T stress_action_p(const T A[3][3], const T g[3][3], T s, T wp, T dv[3][3]) {
  T M[3][3];
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int l = 0; l < 3; l++) M[k][l] = g[0][l] * A[k][0] + g[1][l] * A[k][1] + g[2][l] * A[k][2];
  T S[3][3];
  T s2 = s + s;
  S[0][0] = fma(s2, M[0][0], -wp);
  S[1][1] = fma(s2, M[1][1], -wp);
  S[2][2] = fma(s2, M[2][2], -wp);
  S[0][1] = S[1][0] = s * (M[0][1] + M[1][0]);
  S[0][2] = S[2][0] = s * (M[0][2] + M[2][0]);
  S[1][2] = S[2][1] = s * (M[1][2] + M[2][1]);
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int l = 0; l < 3; l++) dv[k][l] = S[0][l] * A[0][k] + S[1][l] * A[1][k] + S[2][l] * A[2][k];
  return M[0][0] + M[1][1] + M[2][2];
}

// x-transpose accumulation of one quadrature point's dv into the row accumulators
template <int qa, typename T>
__device__ __forceinline__ void acc_row(const T dv[3][3], T P0[3][3], T P1[3][3], T P2[3][3]) {
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      if (nzD(qa, i)) P0[c][i] = fmac(gD(qa, i), dv[0][c], P0[c][i]);
      if (nzB(qa, i)) P1[c][i] = fmac(gB(qa, i), dv[1][c], P1[c][i]);
      if (nzB(qa, i)) P2[c][i] = fmac(gB(qa, i), dv[2][c], P2[c][i]);
    }
}

// x- and y-transpose of one quadrature point's dv straight into the slab
// accumulators (no per-row accumulator kept live):
//   RB[c][j][i] += B(qb,j) D(qa,i) dv[0][c] + D(qb,j) B(qa,i) dv[1][c]
//   RD[c][j][i] += B(qb,j) B(qa,i) dv[2][c]
template <int qa, int qb, typename T>
__device__ __forceinline__ void acc_direct(const T dv[3][3], T RB[3][3][3], T RD[3][3][3]) {
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        if (nzB(qb, j) && nzD(qa, i)) RB[c][j][i] = fmac(gB(qb, j) * gD(qa, i), dv[0][c], RB[c][j][i]);
        if (nzD(qb, j) && nzB(qa, i)) RB[c][j][i] = fmac(gD(qb, j) * gB(qa, i), dv[1][c], RB[c][j][i]);
        if (nzB(qb, j) && nzB(qa, i)) RD[c][j][i] = fmac(gB(qb, j) * gB(qa, i), dv[2][c], RD[c][j][i]);
      }
}

// ---------------------------------------------------------------------------
// Runtime-row variants.  The quadrature row (qb,qc) is a runtime (wave-uniform) loop
// index; its 1D coefficients come from constant memory into scalar registers, so the
// whole element body is one row-sized block of code (a fully unrolled element does
// not fit the instruction cache).  Inside a row everything is unrolled.
__device__ __forceinline__ void row_fence() { __builtin_amdgcn_sched_barrier(0); }

// Asynchronous copy of NB contiguous bytes (a multiple of 16) from global memory into
// LDS with global->LDS loads: lane l moves bytes [16l, 16l+16) of every 1 KiB chunk, and
// no vector register is held while the copy is in flight.  Every lane of the wave must
// take part.  Completion: s_waitcnt vmcnt(0) (WAIT_VM0).
template <typename T>
__device__ __forceinline__ const T *uniform_ptr(const T *p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (const T *)(((uint64_t)hi << 32) | lo);
}
template <int NB>
__device__ __forceinline__ void lds_copy_async(const void *src, void *lds, int lane) {
  static_assert(NB % 16 == 0, "16-byte chunks");
  constexpr int CH = NB / 16, FULL = CH / 64, REM = CH % 64;
  const char *s0 = uniform_ptr((const char *)src);  // scalar base, per-lane byte offset
  // the LDS destination goes to M0: a scalar even where it depends on the wave's index
  typedef __attribute__((address_space(3))) char lchar;
  lchar *d0 = (lchar *)(uintptr_t)__builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(lchar *)lds);
  int off = 16 * lane;
  asm volatile("" : "+v"(off));  // one fresh vector register per copy site
#pragma unroll
  for (int i = 0; i < FULL; i++)
    __builtin_amdgcn_global_load_lds((const void *)(s0 + i * 1024 + off),
                                     (__attribute__((address_space(3))) void *)(d0 + i * 1024), 16, 0, 0);
  if (REM) {
    if (lane < REM)
      __builtin_amdgcn_global_load_lds((const void *)(s0 + FULL * 1024 + off),
                                       (__attribute__((address_space(3))) void *)(d0 + FULL * 1024), 16, 0, 0);
  }
}
// s_waitcnt immediates (gfx9 encoding): all vector-memory operations done (the
// global->LDS copies included) / all LDS operations done
constexpr int WAIT_VM0 = 0x0F70;
constexpr int WAIT_LGKM0 = 0xC07F;

struct Coef1 {
  double b[3], d[3];
};
__constant__ const double c_gB[3][3] = {
    {0.5 * (0.6 + 0.77459666924148337704), 1.0 - 0.6, 0.5 * (0.6 - 0.77459666924148337704)},
    {0.0, 1.0, 0.0},
    {0.5 * (0.6 - 0.77459666924148337704), 1.0 - 0.6, 0.5 * (0.6 + 0.77459666924148337704)}};
__constant__ const double c_gD[3][3] = {{-0.77459666924148337704 - 0.5, 2.0 * 0.77459666924148337704, -0.77459666924148337704 + 0.5},
                                         {-0.5, 0.0, 0.5},
                                         {0.77459666924148337704 - 0.5, -2.0 * 0.77459666924148337704, 0.77459666924148337704 + 0.5}};
__constant__ const double c_gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};

__device__ __forceinline__ Coef1 coef1(int q) {
  Coef1 r;
#pragma unroll
  for (int n = 0; n < 3; n++) {
    r.b[n] = c_gB[q][n];
    r.d[n] = c_gD[q][n];
  }
  return r;
}

// One tile column: the 9 values (j,k) of component c at node column i.
template <typename T>
__device__ __forceinline__ void ld_col(const T *s, int base, int g, T v[9]) {
  const int c = g / 3, i = g % 3;
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int j = 0; j < 3; j++) v[3 * k + j] = tv(s, base, c, i, j, k);
}
// contraction of one column with the row's z (cc) and y (cb) coefficients
template <typename T>
__device__ __forceinline__ void col_contract(const T v[9], const Coef1 &cb, const Coef1 &cc, T &r0, T &r1, T &r2) {
  T tb[3], td[3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    tb[j] = fmac(cc.b[2], v[6 + j], fmac(cc.b[1], v[3 + j], static_cast<T>(cc.b[0]) * v[j]));
    td[j] = fmac(cc.d[2], v[6 + j], fmac(cc.d[1], v[3 + j], static_cast<T>(cc.d[0]) * v[j]));
  }
  r0 = fmac(cb.b[2], tb[2], fmac(cb.b[1], tb[1], static_cast<T>(cb.b[0]) * tb[0]));
  r1 = fmac(cb.d[2], tb[2], fmac(cb.d[1], tb[1], static_cast<T>(cb.d[0]) * tb[0]));
  r2 = fmac(cb.b[2], td[2], fmac(cb.b[1], td[1], static_cast<T>(cb.b[0]) * td[0]));
}
// Row data of one field with the tile reads software-pipelined over two column
// buffers: on entry `first` holds column 0 of field s; column g is contracted from
// `first` (g even) or `second` (g odd) while the next column is read into the other
// buffer.  Column 8 issues the reads of column 0 of field s_next into `second`, so the
// next phase passes the buffers swapped.
template <typename T>
__device__ __forceinline__ void row_data_pipe(const T *s, const T *s_next, int base, const Coef1 &cb,
                                              const Coef1 &cc, T first[9], T second[9], T r0[3][3], T r1[3][3],
                                              T r2[3][3]) {
#pragma unroll
  for (int g = 0; g < 9; g++) {
    T *cur = (g & 1) ? second : first;
    T *nxt = (g & 1) ? first : second;
    if (g < 8)
      ld_col(s, base, g + 1, nxt);
    else
      ld_col(s_next, base, 0, nxt);
    col_contract(cur, cb, cc, r0[g / 3][g % 3], r1[g / 3][g % 3], r2[g / 3][g % 3]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// a finished row (runtime qb,qc) straight into the element output: per (c,j,i) the
// y-transposed pair is formed once and spread over the three z nodes.
template <typename T>
__device__ __forceinline__ void acc_row_elem_rt(const Coef1 &cb, const Coef1 &cc, const T P0[3][3],
                                                const T P1[3][3], const T P2[3][3], T Y[3][3][3][3]) {
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const T tb = fmac(cb.b[j], P0[c][i], static_cast<T>(cb.d[j]) * P1[c][i]);
        const T td = static_cast<T>(cb.b[j]) * P2[c][i];
#pragma unroll
        for (int k = 0; k < 3; k++) Y[c][k][j][i] = fmac(cc.b[k], tb, fmac(cc.d[k], td, Y[c][k][j][i]));
      }
}

// Slab form of the y/z transposes: the rows of one slab (fixed qc) accumulate their y-transposed
// pairs into RB (B-weighted in z) and RD (D-weighted), and acc_slab_rt spreads the slab over the
// three z nodes once:  RB[c][j][i] += B(qb,j) P0 + D(qb,j) P1,  RD += B(qb,j) P2, then
// Y[c][k][j][i] += B(qc,k) RB + D(qc,k) RD.  81 + 54 / 3 operations per row and output instead
// of 243 per row.
template <typename T>
__device__ __forceinline__ void acc_row_slab_rt(const Coef1 &cb, const T P0[3][3], const T P1[3][3],
                                                const T P2[3][3], T RB[3][3][3], T RD[3][3][3]) {
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        RB[c][j][i] = fmac(cb.b[j], P0[c][i], fmac(cb.d[j], P1[c][i], RB[c][j][i]));
        RD[c][j][i] = fmac(cb.b[j], P2[c][i], RD[c][j][i]);
      }
}
template <typename T>
__device__ __forceinline__ void acc_slab_rt(const Coef1 &cc, const T RB[3][3][3], const T RD[3][3][3],
                                            T Y[3][3][3][3]) {
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) Y[c][k][j][i] = fmac(cc.b[k], RB[c][j][i], fmac(cc.d[k], RD[c][j][i], Y[c][k][j][i]));
}

// keep the scheduler from hoisting the next row's LDS reads into this one

// z-transpose of a finished slab into the element output Y[c][k][j][i]
template <int qc, typename T>
__device__ __forceinline__ void acc_elem(const T RB[3][3][3], const T RD[3][3][3], T Y[3][3][3][3]) {
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
          if (nzB(qc, k)) Y[c][k][j][i] = fmac(gB(qc, k), RB[c][j][i], Y[c][k][j][i]);
          if (nzD(qc, k)) Y[c][k][j][i] = fmac(gD(qc, k), RD[c][j][i], Y[c][k][j][i]);
        }
}

template <typename T>
__device__ __forceinline__ void zero33(T a[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) a[i][j] = T(0);
}

}  // namespace ptq2
