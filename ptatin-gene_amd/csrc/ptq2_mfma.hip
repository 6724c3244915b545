// ptq2_mfma.hip -- the FP32 smoother A11 on the matrix cores (MFMA form).
//
// The same operator as k_a11_fp32 (stored factors F = sqrt(w eta / det) adj(J), dv = F^T (M +
// M^T), M = F g), with the reference gradient applied as a dense GEMM: the 81 x 27 operator G
// (rows: quadrature point q and reference direction d, columns: element node) times the nodal
// values of a batch of 16 elements, on v_mfma_f32_16x16x4f32 (16x16 outputs, K = 4 per
// instruction).  The sum-factorized contractions of src/stokes_operators_tensor.c:181-286 are
// replaced by
//   forward   g[(q,d), (c,e)] = sum_n G[(q,d), n] u[n, (c,e)]       (per component c)
//   backward  y[n, (c,e)]     = sum_(q,d) G[(q,d), n] dv[(q,d), (c,e)]
// with the per-point work (M, S, dv) in between on the VALU.  One wave owns a brick of 4x4x4
// elements (the node tiles, factors and owner-writes output of k_a11_fp32) and runs it as four
// batches of 16 elements (one z layer each).
//
// Matrix layouts (16x16x4 f32: A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k, D[i][j] in
// lane j + 16 (i / 4), register i % 4):
//   forward, per quadrature block mb (points 16 mb .. 16 mb + 15), direction d, component c:
//     A = G[(16 mb + i, d), 4 s + k] (K step s: nodes 4 s .. 4 s + 3), B = u[4 s + k][element j],
//     D[i][j] = g of point 16 mb + i, element j: lane j + 16 g holds points 16 mb + 4 g + r (r
//     = register) for every (d, c) tile, so the per-point work is lane-local;
//   backward, per node tile mt (nodes 16 mt .. 16 mt + 15): K steps (mb, r, d) take the points
//     16 mb + 4 k + r (k = 0..3 in the lane groups) with direction d, whose dv sit in register r
//     of the forward tile (mb, d, c) -- the B operand needs no transpose;
//     A = G[(16 mb + 4 k + r, d), 16 mt + i].
// 27 points pad to 32 and 27 nodes to 28 (zero rows of G).  MFMA work per element: 2 x 3 x 7 x
// 3 + 24 x 2 x 3 = 270 instructions of 1024 multiply-adds per 16 elements.
#include "ptq2_device.h"
#include "ptq2_kernels.h"

namespace ptq2 {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

// G[q][d][n] = dN_n / dxi_d at Gauss point q, q = qa + 3 qb + 9 qc, n = i + 3 j + 9 k
struct GTab {
  float v[27 * 3 * 27];
};
constexpr GTab make_gtab() {
  GTab t{};
  for (int q = 0; q < 27; q++)
    for (int n = 0; n < 27; n++) {
      const int qa = q % 3, qb = (q / 3) % 3, qc = q / 9, i = n % 3, j = (n / 3) % 3, k = n / 9;
      t.v[(q * 3 + 0) * 27 + n] = (float)(gD(qa, i) * gB(qb, j) * gB(qc, k));
      t.v[(q * 3 + 1) * 27 + n] = (float)(gB(qa, i) * gD(qb, j) * gB(qc, k));
      t.v[(q * 3 + 2) * 27 + n] = (float)(gB(qa, i) * gB(qb, j) * gD(qc, k));
    }
  return t;
}
__constant__ const GTab c_gtab = make_gtab();

__device__ __forceinline__ float gval(int q, int d, int n) {
  return (q < 27 && n < 27) ? c_gtab.v[(q * 3 + d) * 27 + n] : 0.0f;
}
// LDS offset of element-local node n (component 0) from the element's lane base
__device__ __forceinline__ int node_off(int n) { return xo(n % 3) + TRS * ((n / 3) % 3) + TPS * (n / 9); }

__global__ __launch_bounds__(BEL) void k_a11_fp32_mfma(Box b, const float *__restrict__ u,
                                                       const uint8_t *__restrict__ mask,
                                                       const float *__restrict__ fac, float *__restrict__ y,
                                                       int add_mode, float *__restrict__ facebuf) {
  __shared__ double Ts[3 * TCS];  // output tile (FP64 sums, as in k_a11_fp32)
  __shared__ float Us[3 * TCS];   // velocity tile
  const int lane = threadIdx.x, j = lane & 15, kg = lane >> 4;
  const BrickId br = brick_xcd(b, blockIdx.x);
  const Tile t = tile_of(b, br);
  load_tile(b, t, br, u, mask, Us, lane);
  zero_tile(Ts, lane);
  // the constant A operands (per lane): forward Af[mb][d][s], backward Ab[mb][r][d][mt]
  float Af[2][3][7], Ab[2][4][3][2];
#pragma unroll
  for (int mb = 0; mb < 2; mb++)
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int s = 0; s < 7; s++) Af[mb][d][s] = gval(16 * mb + j, d, 4 * s + kg);
#pragma unroll
  for (int mb = 0; mb < 2; mb++)
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int mt = 0; mt < 2; mt++) Ab[mb][r][d][mt] = gval(16 * mb + 4 * kg + r, d, 16 * mt + j);
  __syncthreads();
  const float *fb = fac + (int64_t)br.lin * 243 * BEL;
#pragma nounroll
  for (int sb = 0; sb < 4; sb++) {
    const int el = 16 * sb + j;  // this lane's element (column j of the batch)
    const int base = lane_base(el);
    const bool active = lane_active(t, el);
    // ---- forward: g = G u
    f4 acc[2][3][3];
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int c = 0; c < 3; c++) acc[mb][d][c] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 7; s++) {
      const int n = 4 * s + kg;
      float bv[3];
#pragma unroll
      for (int c = 0; c < 3; c++) bv[c] = n < 27 ? Us[base + c * TCS + node_off(n)] : 0.0f;
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int mb = 0; mb < 2; mb++)
#pragma unroll
          for (int d = 0; d < 3; d++)
            acc[mb][d][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[mb][d][s], bv[c], acc[mb][d][c], 0, 0, 0);
    }
    // ---- per point: dv = F^T (M + M^T), M = F g; lane (j, kg) holds points 16 mb + 4 kg + r
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int q = 16 * mb + 4 * kg + r;
        float F[9];
        const float *fq = fb + (int64_t)(min(q, 26) * 9) * BEL + el;
#pragma unroll
        for (int m = 0; m < 9; m++) F[m] = q < 27 ? fq[m * BEL] : 0.0f;
        float g[3][3], M[3][3], S[3][3];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
          for (int c = 0; c < 3; c++) g[d][c] = acc[mb][d][c][r];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
          for (int l = 0; l < 3; l++) M[k][l] = g[0][l] * F[3 * k] + g[1][l] * F[3 * k + 1] + g[2][l] * F[3 * k + 2];
        S[0][0] = 2.0f * M[0][0];
        S[1][1] = 2.0f * M[1][1];
        S[2][2] = 2.0f * M[2][2];
        S[0][1] = S[1][0] = M[0][1] + M[1][0];
        S[0][2] = S[2][0] = M[0][2] + M[2][0];
        S[1][2] = S[2][1] = M[1][2] + M[2][1];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
          for (int l = 0; l < 3; l++) acc[mb][k][l][r] = S[0][l] * F[k] + S[1][l] * F[3 + k] + S[2][l] * F[6 + k];
      }
    // ---- backward: y = G^T dv
    f4 yo[2][3];
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
      for (int c = 0; c < 3; c++) yo[mt][c] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
          for (int c = 0; c < 3; c++)
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
              yo[mt][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ab[mb][r][d][mt], acc[mb][d][c][r], yo[mt][c], 0, 0, 0);
    // ---- add the batch's element outputs into the tile (node 16 mt + 4 kg + r of element j)
    if (active) {
#pragma unroll
      for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int n = 16 * mt + 4 * kg + r;
          if (n < 27) {
#pragma unroll
            for (int c = 0; c < 3; c++) lds_add(&Ts[base + c * TCS + node_off(n)], (double)yo[mt][c][r]);
          }
        }
    }
  }
  __syncthreads();
  store_tile_own(b, br, t, Ts, y, facebuf, add_mode != 0, lane);
}

}  // namespace

hipError_t launch_a11_fp32_mfma(const Box &b, const float *u, const uint8_t *mask, const float *fac, float *y,
                                bool add, hipStream_t s, float *facebuf, bool gather) {
  if (!facebuf) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_a11_fp32_mfma, dim3(b.nbx * b.nby * b.nbz), dim3(BEL), 0, s, b, u, mask, fac, y, add ? 1 : 0,
                     facebuf);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !gather) return e;
  return launch_gather_faces(b, (const float *)facebuf, y, s);
}

}  // namespace ptq2
