// ptq2_kernels.hip -- gfx950 kernels of the matrix-free Q2-P1disc Stokes operators.
//
// A11 (viscous block), A12 (gradient), A21 (divergence), the full block and the
// A11 diagonal.  The element algorithm is the sum-factorized 3x3x3 tensor
// contraction of src/stokes_operators_tensor.c:181-286 reorganised for one
// element per lane (see ptq2_device.h for the brick/tile decomposition).
#include "ptq2_device.h"
#include "ptq2_kernels.h"

#include <algorithm>

namespace ptq2 {

// ------------------------------------------------------- pressure basis helper
// P3D_ConstructNi_P1GRel_3D (src/element_utils_q2.c:183-224):
// N^p = [1, (x-xbar)/Lx, (y-ybar)/Ly, (z-zbar)/Lz], xbar the mean of the 27 nodes,
// Lx = x_14 - x_12, Ly = y_16 - y_10, Lz = z_22 - z_4.
struct PBasis {
  double avg[3], inv[3];
};
__device__ __forceinline__ PBasis pbasis_setup(const double *Xs, int base) {
  PBasis pb;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) s += tv(Xs, base, c, i, j, k);
    pb.avg[c] = s * (1.0 / 27.0);
  }
  pb.inv[0] = 1.0 / (tv(Xs, base, 0, 2, 1, 1) - tv(Xs, base, 0, 0, 1, 1));
  pb.inv[1] = 1.0 / (tv(Xs, base, 1, 1, 2, 1) - tv(Xs, base, 1, 1, 0, 1));
  pb.inv[2] = 1.0 / (tv(Xs, base, 2, 1, 1, 2) - tv(Xs, base, 2, 1, 1, 0));
  return pb;
}


// ----------------------------------------------------------------- A11 FP64

// geometry of quadrature point qa of the current row: adjugate and s = w eta / det
template <int qa>
__device__ __forceinline__ double geo_qp(const double xr0[3][3], const double xr1[3][3], const double xr2[3][3],
                                         double wbc, double eta, double A[3][3], double &s) {
  double J[3][3];
  qp_grad<qa>(xr0, xr1, xr2, J);
  double det = adjugate(J, A);
  s = (gw1(qa) * wbc * eta) * rcp64(det);
  return det;
}

// Pressure-side state of the full block (BLK): the element's P1GRel basis set-up and
// pressure, the continuity accumulators and, for the residual (fp != nullptr), the 27
// quadrature values of the continuity source.
struct BlockState {
  PBasis pb;
  double pe[4];
  double Yp[4];
  const double *fp;
};

// pressure basis values N^p_1..3 at quadrature point qa of the row (x from the B-B row data)
template <int qa>
__device__ __forceinline__ void pbasis_qp(const double xr0[3][3], const PBasis &pb, double np[3]) {
  double x[3];
  qp_value<qa>(xr0, x);
#pragma unroll
  for (int d = 0; d < 3; d++) np[d] = (x[d] - pb.avg[d]) * pb.inv[d];
}

// one quadrature point of the velocity phase
template <int qa, bool BLK>
__device__ __forceinline__ void a11_qp64(const double ur0[3][3], const double ur1[3][3], const double ur2[3][3],
                                         const double A[3][3], double s, double wbc, const double np[3],
                                         double wdet, int q, BlockState *bs, double P0[3][3], double P1[3][3],
                                         double P2[3][3]) {
  double g[3][3], dv[3][3];
  qp_grad<qa>(ur0, ur1, ur2, g);
  if constexpr (BLK) {
    const double w = gw1(qa) * wbc;
    const double pq = bs->pe[0] + bs->pe[1] * np[0] + bs->pe[2] * np[1] + bs->pe[3] * np[2];
    const double f = -w * stress_action_p(A, g, s, w * pq, dv);
    double fs = 0.0;
    if (bs->fp) fs = -wdet * bs->fp[q];
    bs->Yp[0] += f + fs;
#pragma unroll
    for (int d = 0; d < 3; d++) bs->Yp[1 + d] += (f + fs) * np[d];
  } else {
    stress_action(A, g, s, dv);
  }
  acc_row<qa>(dv, P0, P1, P2);
}

// one quadrature row (runtime qb,qc) of the FP64 A11 element: geometry of its three
// points from the coordinate tile, velocity gradients from the velocity tile, the
// stress action and the x-transpose into P, then the y- and z-transposes into Y.
// On entry ca holds column 0 of the coordinate tile; on exit it holds it again (for
// the next row), its reads in flight behind the stress work.  BLK adds the pressure and
// continuity terms of the full block.
template <bool BLK, bool SLAB = false>
__device__ __forceinline__ void a11_row64(const double *Xs, const double *Us, int base, const Coef1 &cb,
                                          const Coef1 &cc, double wbc, const double e3[3], double ca[9],
                                          double cb9[9], double Y[3][3][3][3], int row, BlockState *bs,
                                          double (*RB)[3][3] = nullptr, double (*RD)[3][3] = nullptr) {
  double A0[3][3], A1[3][3], A2[3][3], s0, s1, s2;
  double np[3][3], wd[3] = {0.0, 0.0, 0.0};
  {
    double xr0[3][3], xr1[3][3], xr2[3][3];
    row_data_pipe(Xs, Us, base, cb, cc, ca, cb9, xr0, xr1, xr2);
    const double d0 = geo_qp<0>(xr0, xr1, xr2, wbc, e3[0], A0, s0);
    const double d1 = geo_qp<1>(xr0, xr1, xr2, wbc, e3[1], A1, s1);
    const double d2 = geo_qp<2>(xr0, xr1, xr2, wbc, e3[2], A2, s2);
    if constexpr (BLK) {
      pbasis_qp<0>(xr0, bs->pb, np[0]);
      pbasis_qp<1>(xr0, bs->pb, np[1]);
      pbasis_qp<2>(xr0, bs->pb, np[2]);
      if (bs->fp) {
        wd[0] = gw1(0) * wbc * d0;
        wd[1] = gw1(1) * wbc * d1;
        wd[2] = gw1(2) * wbc * d2;
      }
    } else {
      (void)d0; (void)d1; (void)d2;
    }
  }
  row_fence();
  double ur0[3][3], ur1[3][3], ur2[3][3];
  row_data_pipe(Us, Xs, base, cb, cc, cb9, ca, ur0, ur1, ur2);
  double P0[3][3], P1[3][3], P2[3][3];
  zero33(P0); zero33(P1); zero33(P2);
  a11_qp64<0, BLK>(ur0, ur1, ur2, A0, s0, wbc, np[0], wd[0], 3 * row, bs, P0, P1, P2);
  a11_qp64<1, BLK>(ur0, ur1, ur2, A1, s1, wbc, np[1], wd[1], 3 * row + 1, bs, P0, P1, P2);
  a11_qp64<2, BLK>(ur0, ur1, ur2, A2, s2, wbc, np[2], wd[2], 3 * row + 2, bs, P0, P1, P2);
  row_fence();
  if constexpr (SLAB)
    acc_row_slab_rt(cb, P0, P1, P2, RB, RD);
  else
    acc_row_elem_rt(cb, cc, P0, P1, P2, Y);
  row_fence();
}

// add the element's 81 outputs into the LDS output tile (of the element's type or wider)
template <typename S, typename T>
__device__ __forceinline__ void emit_element(S *Ys, int base, const T Y[3][3][3][3]) {
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) lds_add(&Ys[base + toff(c, i, j, k)], S(Y[c][k][j][i]));
}

__device__ __forceinline__ void load9(const double *__restrict__ ep, int qc, double *e9) {
#pragma unroll
  for (int q = 0; q < 9; q++) e9[q] = ep[(9 * qc + q) * BEL];
}

__device__ __forceinline__ void load3(const double *__restrict__ ep, int row, double *e3) {
#pragma unroll
  for (int q = 0; q < 3; q++) e3[q] = ep[(3 * row + q) * BEL];
}

// BLK = false: the A11 apply.  BLK = true: the full block [A11 A12; A21 0] in the same single
// element pass (MFStokesWrapper_A_AVX, src/stokes_operators_avx.c:439-636): y gets A11 u + A12 p,
// yp gets A21 u (and, with fp, minus the continuity source: FormFunctionLocal_P's
// -int N^p fp, src/stokes_form_function.c:478-580).
// CH: store with the fused Chebyshev update of co (store_tile_own<true>).
// The brick br of box b (one wave); k_a11_fp64_t and the batched k_a11_fp64_batch resolve br.
template <bool BLK, bool CH>
__device__ __forceinline__ void a11_fp64_brick(const Box &b, const BrickId &br, int color,
                                               const double *__restrict__ coords, const double *__restrict__ u,
                                               const uint8_t *__restrict__ mask, const double *__restrict__ eta_bm,
                                               double *__restrict__ y, int add_mode, double *__restrict__ facebuf,
                                               const double *__restrict__ p, double *__restrict__ yp,
                                               const double *__restrict__ fp, const ChebOut<double> &co) {
  __shared__ double Xs[3 * TCS];
  __shared__ double Us[3 * TCS];
  const int lane = threadIdx.x;
  const Tile t = tile_of(b, br);
  const double *ep = eta_bm + (int64_t)br.lin * 27 * BEL + lane;
  double en[3];
  load3(ep, 0, en);
  load_tile2(b, t, br, coords, Xs, u, mask, Us, lane);
  __syncthreads();
  const int base = lane_base(lane);
  const bool active = lane_active(t, lane);
  double Y[3][3][3][3];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) Y[c][k][j][i] = 0.0;
  BlockState bs;
  int64_t e = 0;
  if constexpr (BLK) {
    bs.Yp[0] = bs.Yp[1] = bs.Yp[2] = bs.Yp[3] = 0.0;
    bs.pe[0] = bs.pe[1] = bs.pe[2] = bs.pe[3] = 0.0;
    bs.fp = nullptr;
    if (active) {
      e = lane_element(b, t, lane);
      bs.pb = pbasis_setup(Xs, base);
#pragma unroll
      for (int i = 0; i < 4; i++) bs.pe[i] = p[4 * e + i];
      if (fp) bs.fp = fp + 27 * e;
    }
  }
  if (active) {
    double ca[9], cb9[9];
    ld_col(Xs, base, 0, ca);
    if constexpr (!BLK) {
      // slab form of the y/z transposes (the full block keeps the per-row form: with its
      // pressure state the slab sums would spill)
#pragma nounroll
      for (int qc = 0; qc < 3; qc++) {
        double RB[3][3][3], RD[3][3][3];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
          for (int j = 0; j < 3; j++)
#pragma unroll
            for (int i = 0; i < 3; i++) RB[c][j][i] = RD[c][j][i] = 0.0;
#pragma nounroll
        for (int qb = 0; qb < 3; qb++) {
          const int row = qb + 3 * qc;
          double e3[3] = {en[0], en[1], en[2]};
          if (row < 8) load3(ep, row + 1, en);
          a11_row64<BLK, true>(Xs, Us, base, coef1(qb), coef1(qc), c_gw[qb] * c_gw[qc], e3, ca, cb9, Y, row, &bs,
                               RB, RD);
        }
        acc_slab_rt(coef1(qc), RB, RD, Y);
        row_fence();
      }
    } else {
#pragma nounroll
      for (int row = 0; row < 9; row++) {
        const int qb = row % 3, qc = row / 3;
        double e3[3] = {en[0], en[1], en[2]};
        if (row < 8) load3(ep, row + 1, en);
        a11_row64<BLK>(Xs, Us, base, coef1(qb), coef1(qc), c_gw[qb] * c_gw[qc], e3, ca, cb9, Y, row, &bs);
      }
    }
    if constexpr (BLK) {
      double *o = yp + 4 * e;
#pragma unroll
      for (int i = 0; i < 4; i++) o[i] = add_mode ? o[i] + bs.Yp[i] : bs.Yp[i];
    }
  }
  __syncthreads();
  zero_tile(Xs, lane);
  __syncthreads();
  if (active) emit_element(Xs, base, Y);
  __syncthreads();
  if (color >= 0)
    store_tile(b, br, t, Xs, y, add_mode != 0, lane);
  else if constexpr (CH)
    store_tile_own<true>(b, br, t, Xs, y, facebuf, false, lane, &co);
  else
    store_tile_own(b, br, t, Xs, y, facebuf, add_mode != 0, lane);
}

template <bool BLK, bool CH>
__global__ __launch_bounds__(BEL) void k_a11_fp64_t(Box b, int color, const double *__restrict__ coords,
                                                    const double *__restrict__ u, const uint8_t *__restrict__ mask,
                                                    const double *__restrict__ eta_bm, double *__restrict__ y,
                                                    int add_mode, double *__restrict__ facebuf,
                                                    const double *__restrict__ p, double *__restrict__ yp,
                                                    const double *__restrict__ fp, ChebOut<double> co) {
  const BrickId br = color >= 0 ? brick_of(b, color, blockIdx.x) : brick_xcd(b, blockIdx.x);
  a11_fp64_brick<BLK, CH>(b, br, color, coords, u, mask, eta_bm, y, add_mode, facebuf, p, yp, fp, co);
}

// Several problems in one launch: block i belongs to the problem p with start[p] <= i <
// start[p + 1] (starts are multiples of 8, so i and the local block index fall on the same
// XCD and brick_xcd keeps its L2 grouping); the blocks past a problem's bricks return.
__device__ __forceinline__ int batch_problem(const int *start, int n, int bid) {
  int p = 0;
#pragma unroll
  for (int i = 1; i < BATCH_MAX; i++) p += (i < n && bid >= start[i]) ? 1 : 0;
  return p;
}

__global__ __launch_bounds__(BEL) void k_a11_fp64_batch(BatchArgs<double> a) {
  const int p = batch_problem(a.start, a.n, blockIdx.x);
  const Box b = a.box[p];
  const int local = (int)blockIdx.x - a.start[p];
  if (local >= b.nbx * b.nby * b.nbz) return;
  a11_fp64_brick<false, false>(b, brick_xcd(b, local), -1, a.coords[p], a.u[p], a.mask[p], a.fac[p], a.y[p], a.add,
                               a.face[p], nullptr, nullptr, nullptr, ChebOut<double>{});
}

// ----------------------------------------------------------------- A11 FP32
// Geometric factors are precomputed in FP64 (k_build_fac32) and stored rounded
// to FP32 as F = sqrt(w eta / det) * adj(J), 9 per quadrature point, brick-major
// [brick][q][m][lane] so that every load instruction is one contiguous 256-byte
// segment.  Then dv = F^T (M + M^T) with M = F g (diagonal doubled), all FP32.

// one quadrature point of the FP32 A11 row: F = sqrt(w eta / det) adj(J) (9 floats)
template <int qa>
__device__ __forceinline__ void a11_qp32(const float ur0[3][3], const float ur1[3][3], const float ur2[3][3],
                                         const float F9[9], float P0[3][3], float P1[3][3], float P2[3][3]) {
  float g[3][3], dv[3][3];
  qp_grad<qa>(ur0, ur1, ur2, g);
  float M[3][3];
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int l = 0; l < 3; l++) M[k][l] = g[0][l] * F9[3 * k] + g[1][l] * F9[3 * k + 1] + g[2][l] * F9[3 * k + 2];
  float S[3][3];
  S[0][0] = 2.0f * M[0][0];
  S[1][1] = 2.0f * M[1][1];
  S[2][2] = 2.0f * M[2][2];
  S[0][1] = S[1][0] = M[0][1] + M[1][0];
  S[0][2] = S[2][0] = M[0][2] + M[2][0];
  S[1][2] = S[2][1] = M[1][2] + M[2][1];
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int l = 0; l < 3; l++) dv[k][l] = S[0][l] * F9[k] + S[1][l] * F9[3 + k] + S[2][l] * F9[6 + k];
  acc_row<qa>(dv, P0, P1, P2);
}

// z-contraction of the velocity tile for one slab (runtime qc): for every column
// (c,i,j) the B- and D-weighted sums over the three z nodes.  The three rows of the slab
// reuse them, so the tile is read once per slab instead of once per row.
__device__ __forceinline__ void slab_z32(const float *Us, int base, const Coef1 &cc, float ZB[3][3][3],
                                         float ZD[3][3][3]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    float v[3][3][3];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) v[k][j][i] = tv(Us, base, c, i, j, k);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        ZB[c][i][j] = fmac(cc.b[2], v[2][j][i], fmac(cc.b[1], v[1][j][i], static_cast<float>(cc.b[0]) * v[0][j][i]));
        ZD[c][i][j] = fmac(cc.d[2], v[2][j][i], fmac(cc.d[1], v[1][j][i], static_cast<float>(cc.d[0]) * v[0][j][i]));
      }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// FP32 factor row of one quadrature row for the 64 lanes: [qa*9+m][lane]
constexpr int FROW = 27 * BEL;

// Quadrature rows row0 .. row1-1 of the FP32 element, accumulated into Y.  On entry the factors
// of row0 are on their way into Fs (lds_copy_async from fsrc + row0 * FROW); each row's
// successor is copied in while it computes.  row0 starts a slab (row0 % 3 == 0).
__device__ __forceinline__ void fp32_rows(const float *Us, float *Fs, const float *fsrc, int base, int lane, int row0,
                                          int row1, float Y[3][3][3][3]) {
  float ZB[3][3][3], ZD[3][3][3];
#pragma nounroll
  for (int row = row0; row < row1; row++) {
    const int qb = row % 3, qc = row / 3;
    const Coef1 cb = coef1(qb), cc = coef1(qc);
    __builtin_amdgcn_s_waitcnt(WAIT_VM0);
    float F0[9], F12[18];
#pragma unroll
    for (int n = 0; n < 9; n++) F0[n] = Fs[n * BEL + lane];
#pragma unroll
    for (int n = 0; n < 18; n++) F12[n] = Fs[(9 + n) * BEL + lane];
    if (row + 1 < row1) {
      __builtin_amdgcn_s_waitcnt(WAIT_LGKM0);
      lds_copy_async<FROW * 4>(fsrc + (row + 1) * FROW, Fs, lane);
    }
    if (qb == 0) slab_z32(Us, base, cc, ZB, ZD);
    float ur0[3][3], ur1[3][3], ur2[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        ur0[c][i] = fmac(cb.b[2], ZB[c][i][2], fmac(cb.b[1], ZB[c][i][1], static_cast<float>(cb.b[0]) * ZB[c][i][0]));
        ur1[c][i] = fmac(cb.d[2], ZB[c][i][2], fmac(cb.d[1], ZB[c][i][1], static_cast<float>(cb.d[0]) * ZB[c][i][0]));
        ur2[c][i] = fmac(cb.b[2], ZD[c][i][2], fmac(cb.b[1], ZD[c][i][1], static_cast<float>(cb.b[0]) * ZD[c][i][0]));
      }
    float P0[3][3], P1[3][3], P2[3][3];
    zero33(P0); zero33(P1); zero33(P2);
    a11_qp32<0>(ur0, ur1, ur2, F0, P0, P1, P2);
    a11_qp32<1>(ur0, ur1, ur2, F12, P0, P1, P2);
    a11_qp32<2>(ur0, ur1, ur2, F12 + 9, P0, P1, P2);
    row_fence();
    acc_row_elem_rt(cb, cc, P0, P1, P2, Y);
    row_fence();
  }
}

template <bool CH>
__device__ __forceinline__ void a11_fp32_brick(const Box &b, const BrickId &br, int color, const float *__restrict__ u,
                                               const uint8_t *__restrict__ mask, const float *__restrict__ fac,
                                               float *__restrict__ y, int add_mode, float *__restrict__ facebuf,
                                               const ChebOut<float> &co) {
  // LDS: the FP32 velocity tile and one quadrature row of factors ([qa*9+m][lane]) during
  // the rows, then the output tile in FP64 (LDS float atomics are many times slower than
  // ds_add_f64 on gfx950, and the wider sum costs nothing in the result)
  __shared__ double Ts[3 * TCS];
  float *Us = reinterpret_cast<float *>(Ts);
  float *Fs = Us + 3 * TCS;
  static_assert(sizeof(float) * (3 * TCS + FROW) <= sizeof(Ts), "velocity tile + factor row fit the output tile");
  const int lane = threadIdx.x;
  const Tile t = tile_of(b, br);
  const float *fsrc = fac + (int64_t)br.lin * 243 * BEL;
  lds_copy_async<FROW * 4>(fsrc, Fs, lane);
  load_tile(b, t, br, u, mask, Us, lane);
  __syncthreads();
  const int base = lane_base(lane);
  const bool active = lane_active(t, lane);
  float Y[3][3][3][3];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) Y[c][k][j][i] = 0.0f;
  // Row r+1's factors travel HBM -> LDS while row r computes.  Every lane runs the rows
  // (the copies need the whole wave); the inactive lanes of a partial brick compute on
  // tile and factor slots that no element owns and their outputs are discarded.
  fp32_rows(Us, Fs, fsrc, base, lane, 0, 9, Y);
  __syncthreads();
  zero_tile(Ts, lane);
  __syncthreads();
  if (active) emit_element(Ts, base, Y);
  __syncthreads();
  if (color >= 0)
    store_tile(b, br, t, Ts, y, add_mode != 0, lane);
  else if constexpr (CH)
    store_tile_own<true>(b, br, t, Ts, y, facebuf, false, lane, &co);
  else
    store_tile_own(b, br, t, Ts, y, facebuf, add_mode != 0, lane);
}

template <bool CH>
__global__ __launch_bounds__(BEL) void k_a11_fp32(Box b, int color, const float *__restrict__ u,
                                                  const uint8_t *__restrict__ mask, const float *__restrict__ fac,
                                                  float *__restrict__ y, int add_mode, float *__restrict__ facebuf,
                                                  ChebOut<float> co) {
  const BrickId br = color >= 0 ? brick_of(b, color, blockIdx.x) : brick_xcd(b, blockIdx.x);
  a11_fp32_brick<CH>(b, br, color, u, mask, fac, y, add_mode, facebuf, co);
}

__global__ __launch_bounds__(BEL) void k_a11_fp32_batch(BatchArgs<float> a) {
  const int p = batch_problem(a.start, a.n, blockIdx.x);
  const Box b = a.box[p];
  const int local = (int)blockIdx.x - a.start[p];
  if (local >= b.nbx * b.nby * b.nbz) return;
  a11_fp32_brick<false>(b, brick_xcd(b, local), -1, a.u[p], a.mask[p], a.fac[p], a.y[p], a.add, a.face[p],
                        ChebOut<float>{});
}

// ------------------------------------------------ row-split variants (small levels)
// Three waves per brick: wave w runs the slab of quadrature rows 3w .. 3w+2 (qc = w) into its own
// partial element outputs, and the partials go into the output tile wave after wave (w = 0, 1,
// 2), so the result does not depend on the waves' timing.  A brick's rows take a third of the
// time, which pays where a level has too few bricks to fill the GPU; the sum is taken in
// another order than the one-wave kernels' (rounding differs in the last bits).  One-launch
// output path only (facebuf).
__device__ __forceinline__ int split_wave() { return __builtin_amdgcn_readfirstlane((int)threadIdx.x / BEL); }

__global__ __launch_bounds__(3 * BEL) void k_a11_fp64_split(Box b, const double *__restrict__ coords,
                                                            const double *__restrict__ u,
                                                            const uint8_t *__restrict__ mask,
                                                            const double *__restrict__ eta_bm, double *__restrict__ y,
                                                            int add_mode, double *__restrict__ facebuf) {
  __shared__ double Xs[3 * TCS];
  __shared__ double Us[3 * TCS];
  const int w = split_wave(), lane = (int)threadIdx.x - BEL * w;
  const BrickId br = brick_xcd(b, blockIdx.x);
  const Tile t = tile_of(b, br);
  const double *ep = eta_bm + (int64_t)br.lin * 27 * BEL + lane;
  double en[3];
  load3(ep, 3 * w, en);
  if (w == 0) load_tile2(b, t, br, coords, Xs, u, mask, Us, lane);
  __syncthreads();
  const int base = lane_base(lane);
  const bool active = lane_active(t, lane);
  double Y[3][3][3][3];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) Y[c][k][j][i] = 0.0;
  BlockState bs;
  if (active) {
    double ca[9], cb9[9], RB[3][3][3], RD[3][3][3];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) RB[c][j][i] = RD[c][j][i] = 0.0;
    ld_col(Xs, base, 0, ca);
#pragma nounroll
    for (int row = 3 * w; row < 3 * w + 3; row++) {
      const int qb = row % 3;
      double e3[3] = {en[0], en[1], en[2]};
      if (row + 1 < 3 * w + 3) load3(ep, row + 1, en);
      a11_row64<false, true>(Xs, Us, base, coef1(qb), coef1(w), c_gw[qb] * c_gw[w], e3, ca, cb9, Y, row, &bs, RB,
                             RD);
    }
    acc_slab_rt(coef1(w), RB, RD, Y);  // the wave's slab (qc = w) into its outputs
  }
  __syncthreads();
  if (w == 0) zero_tile(Xs, lane);
  __syncthreads();
  for (int k = 0; k < 3; k++) {
    if (w == k && active) emit_element(Xs, base, Y);
    __syncthreads();
  }
  if (w == 0) store_tile_own(b, br, t, Xs, y, facebuf, add_mode != 0, lane);
}

__global__ __launch_bounds__(3 * BEL) void k_a11_fp32_split(Box b, const float *__restrict__ u,
                                                            const uint8_t *__restrict__ mask,
                                                            const float *__restrict__ fac, float *__restrict__ y,
                                                            int add_mode, float *__restrict__ facebuf) {
  __shared__ double Ts[3 * TCS];
  __shared__ float Us[3 * TCS];
  __shared__ float Fs[3][FROW];
  const int w = split_wave(), lane = (int)threadIdx.x - BEL * w;
  const BrickId br = brick_xcd(b, blockIdx.x);
  const Tile t = tile_of(b, br);
  const float *fsrc = fac + (int64_t)br.lin * 243 * BEL;
  lds_copy_async<FROW * 4>(fsrc + 3 * w * FROW, Fs[w], lane);
  if (w == 0) load_tile(b, t, br, u, mask, Us, lane);
  if (w == 1) zero_tile(Ts, lane);
  __syncthreads();
  const int base = lane_base(lane);
  const bool active = lane_active(t, lane);
  float Y[3][3][3][3];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) Y[c][k][j][i] = 0.0f;
  fp32_rows(Us, Fs[w], fsrc, base, lane, 3 * w, 3 * w + 3, Y);
  __syncthreads();
  for (int k = 0; k < 3; k++) {
    if (w == k && active) emit_element(Ts, base, Y);
    __syncthreads();
  }
  if (w == 0) store_tile_own(b, br, t, Ts, y, facebuf, add_mode != 0, lane);
}

// ------------------------------------------------------ A11 FP64, stored geometry
// The same operator with the per-point geometry G = sqrt(w/det) adj(J) resident in HBM
// (FP64, built once per mesh by k_build_geo64) and eta applied at run time:
// M = G grad u, S = eta (M + M^T), dv = G^T S equals w det J^-T (2 eta eps(u)).  Trades
// 1944 B/element of reads for the coordinate tile and the Jacobian work of k_a11_fp64.
//   geo [brick][q][9][lane] doubles
template <int qa>
__device__ __forceinline__ void a11_qp64g(const double ur0[3][3], const double ur1[3][3], const double ur2[3][3],
                                          const double G9[9], double eta, double P0[3][3], double P1[3][3],
                                          double P2[3][3]) {
  double g[3][3], dv[3][3];
  qp_grad<qa>(ur0, ur1, ur2, g);
  double M[3][3];
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int l = 0; l < 3; l++) M[k][l] = g[0][l] * G9[3 * k] + g[1][l] * G9[3 * k + 1] + g[2][l] * G9[3 * k + 2];
  double S[3][3];
  const double e2 = eta + eta;
  S[0][0] = e2 * M[0][0];
  S[1][1] = e2 * M[1][1];
  S[2][2] = e2 * M[2][2];
  S[0][1] = S[1][0] = eta * (M[0][1] + M[1][0]);
  S[0][2] = S[2][0] = eta * (M[0][2] + M[2][0]);
  S[1][2] = S[2][1] = eta * (M[1][2] + M[2][1]);
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int l = 0; l < 3; l++) dv[k][l] = S[0][l] * G9[k] + S[1][l] * G9[3 + k] + S[2][l] * G9[6 + k];
  acc_row<qa>(dv, P0, P1, P2);
}

__global__ __launch_bounds__(BEL) void k_a11_fp64g(Box b, int color, const double *__restrict__ u,
                                                   const uint8_t *__restrict__ mask, const double *__restrict__ geo,
                                                   const double *__restrict__ eta_bm, double *__restrict__ y,
                                                   int add_mode, double *__restrict__ facebuf) {
  __shared__ double Us[3 * TCS];
  const int lane = threadIdx.x;
  const BrickId br = color >= 0 ? brick_of(b, color, blockIdx.x) : brick_xcd(b, blockIdx.x);
  const Tile t = tile_of(b, br);
  const double *gbase = geo + (int64_t)br.lin * 243 * BEL + lane;
  const double *ep = eta_bm + (int64_t)br.lin * 27 * BEL + lane;
  load_tile(b, t, br, u, mask, Us, lane);
  __syncthreads();
  const int base = lane_base(lane);
  const bool active = lane_active(t, lane);
  double Y[3][3][3][3];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) Y[c][k][j][i] = 0.0;
  if (active) {
    double ca[9], cb9[9];
    ld_col(Us, base, 0, ca);
#pragma nounroll
    for (int row = 0; row < 9; row++) {
      const int qb = row % 3, qc = row / 3;
      const Coef1 cb = coef1(qb), cc = coef1(qc);
      double G[27], e3[3];
#pragma unroll
      for (int n = 0; n < 27; n++) G[n] = gbase[(int64_t)(27 * row + n) * BEL];
#pragma unroll
      for (int q = 0; q < 3; q++) e3[q] = ep[(3 * row + q) * BEL];
      double ur0[3][3], ur1[3][3], ur2[3][3];
      row_data_pipe(Us, Us, base, cb, cc, ca, cb9, ur0, ur1, ur2);
#pragma unroll
      for (int n = 0; n < 9; n++) ca[n] = cb9[n];
      double P0[3][3], P1[3][3], P2[3][3];
      zero33(P0); zero33(P1); zero33(P2);
      a11_qp64g<0>(ur0, ur1, ur2, G, e3[0], P0, P1, P2);
      a11_qp64g<1>(ur0, ur1, ur2, G + 9, e3[1], P0, P1, P2);
      a11_qp64g<2>(ur0, ur1, ur2, G + 18, e3[2], P0, P1, P2);
      row_fence();
      acc_row_elem_rt(cb, cc, P0, P1, P2, Y);
      row_fence();
    }
  }
  __syncthreads();
  zero_tile(Us, lane);
  __syncthreads();
  if (active) emit_element(Us, base, Y);
  __syncthreads();
  if (color >= 0)
    store_tile(b, br, t, Us, y, add_mode != 0, lane);
  else
    store_tile_own(b, br, t, Us, y, facebuf, add_mode != 0, lane);
}

// G = sqrt(w/det) adj(J) per point (FP64), for k_a11_fp64g
template <int qb, int qc>
__device__ __forceinline__ void geo_row(const double *Xs, int base, double *__restrict__ gp) {
  double xr0[3][3], xr1[3][3], xr2[3][3];
  row_data<qb, qc>(Xs, base, xr0, xr1, xr2);
#pragma unroll
  for (int qa = 0; qa < 3; qa++) {
    double J[3][3], A[3][3];
    switch (qa) {
      case 0: qp_grad<0>(xr0, xr1, xr2, J); break;
      case 1: qp_grad<1>(xr0, xr1, xr2, J); break;
      default: qp_grad<2>(xr0, xr1, xr2, J); break;
    }
    const double det = adjugate(J, A);
    const int q = qa + 3 * qb + 9 * qc;
    const double w = gw1(qa) * gw1(qb) * gw1(qc);
    const double sc = det > 0.0 ? sqrt(w / det) : 0.0;
#pragma unroll
    for (int m = 0; m < 9; m++) gp[(q * 9 + m) * BEL] = sc * A[m / 3][m % 3];
  }
  row_fence();
}

__global__ __launch_bounds__(BEL) void k_build_geo64(Box b, const double *__restrict__ coords, double *__restrict__ geo) {
  __shared__ double Xs[3 * TCS];
  const int lane = threadIdx.x;
  const BrickId br = brick_lin(b, blockIdx.x);
  const Tile t = tile_of(b, br);
  load_tile(b, t, br, coords, (const uint8_t *)nullptr, Xs, lane);
  __syncthreads();
  if (!lane_active(t, lane)) return;
  const int base = lane_base(lane);
  double *gp = geo + (int64_t)br.lin * 243 * BEL + lane;
  geo_row<0, 0>(Xs, base, gp); geo_row<1, 0>(Xs, base, gp); geo_row<2, 0>(Xs, base, gp);
  geo_row<0, 1>(Xs, base, gp); geo_row<1, 1>(Xs, base, gp); geo_row<2, 1>(Xs, base, gp);
  geo_row<0, 2>(Xs, base, gp); geo_row<1, 2>(Xs, base, gp); geo_row<2, 2>(Xs, base, gp);
}

// ------------------------------------------------- FP32 full block [A11 A12; A21 0]
// One pass: the A11 point action of k_a11_fp32 plus, per point, p_q = pe . N^p and
//   dv[k][l] -= p_q w adj(J)[l][k]            (A12, QuadratureAction_A12_AVX)
//   yp_i     -= N^p_i w trace(adj(J) grad u)  (A21, MatMultMF_Stokes_MixedFEM3d_A21)
// with w adj(J) = facb0 F taken from the FP32 factors (MFStokesWrapper_A_AVX,
// src/stokes_operators_avx.c:439-636, in FP32 with FP32-stored geometry).
template <int qa>
__device__ __forceinline__ void stokes_qp32(const float ur0[3][3], const float ur1[3][3], const float ur2[3][3],
                                            const float F9[9], const float B4[4], const float pe[4], float P0[3][3],
                                            float P1[3][3], float P2[3][3], float Yp[4]) {
  float g[3][3], dv[3][3];
  qp_grad<qa>(ur0, ur1, ur2, g);
  float M[3][3];
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int l = 0; l < 3; l++) M[k][l] = g[0][l] * F9[3 * k] + g[1][l] * F9[3 * k + 1] + g[2][l] * F9[3 * k + 2];
  float S[3][3];
  S[0][0] = 2.0f * M[0][0];
  S[1][1] = 2.0f * M[1][1];
  S[2][2] = 2.0f * M[2][2];
  S[0][1] = S[1][0] = M[0][1] + M[1][0];
  S[0][2] = S[2][0] = M[0][2] + M[2][0];
  S[1][2] = S[2][1] = M[1][2] + M[2][1];
  const float pq = pe[0] + pe[1] * B4[1] + pe[2] * B4[2] + pe[3] * B4[3];
  const float fp = -pq * B4[0];
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int l = 0; l < 3; l++)
      dv[k][l] = S[0][l] * F9[k] + S[1][l] * F9[3 + k] + S[2][l] * F9[6 + k] + fp * F9[3 * l + k];
  const float f = -B4[0] * (M[0][0] + M[1][1] + M[2][2]);
  Yp[0] += f;
  Yp[1] += f * B4[1];
  Yp[2] += f * B4[2];
  Yp[3] += f * B4[3];
  acc_row<qa>(dv, P0, P1, P2);
}

__global__ __launch_bounds__(BEL) void k_stokes_fp32(Box b, int color, const float *__restrict__ u,
                                                     const float *__restrict__ p, const uint8_t *__restrict__ mask,
                                                     const float *__restrict__ fac, const float *__restrict__ facb,
                                                     float *__restrict__ yu, float *__restrict__ yp, int add_mode,
                                                     float *__restrict__ facebuf) {
  // LDS: the FP32 velocity tile and one quadrature row of A11 factors ([qa*9+m][lane]) and
  // block factors ([qa*4+n][lane]) during the rows, then the FP64 output tile
  constexpr int FROW = 27 * BEL, BROW = 12 * BEL;
  constexpr int NTS = (sizeof(float) * (3 * TCS + FROW + BROW) + sizeof(double) - 1) / sizeof(double);
  static_assert(NTS >= 3 * TCS, "the output tile fits");
  __shared__ double Ts[NTS];
  float *Us = reinterpret_cast<float *>(Ts);
  float *Fs = Us + 3 * TCS, *Bs = Fs + FROW;
  const int lane = threadIdx.x;
  const BrickId br = color >= 0 ? brick_of(b, color, blockIdx.x) : brick_xcd(b, blockIdx.x);
  const Tile t = tile_of(b, br);
  const float *fsrc = fac + (int64_t)br.lin * 243 * BEL;
  const float *bsrc = facb + (int64_t)br.lin * 108 * BEL;
  lds_copy_async<FROW * 4>(fsrc, Fs, lane);
  lds_copy_async<BROW * 4>(bsrc, Bs, lane);
  load_tile(b, t, br, u, mask, Us, lane);
  __syncthreads();
  const int base = lane_base(lane);
  const bool active = lane_active(t, lane);
  const int64_t e = lane_element(b, t, lane);
  float Y[3][3][3][3];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) Y[c][k][j][i] = 0.0f;
  float Yp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  // every lane runs the rows (the copies need the whole wave); inactive lanes of a partial
  // brick read no pressure and their outputs are discarded
  {
    float pe[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (active) {
#pragma unroll
      for (int n = 0; n < 4; n++) pe[n] = p[4 * e + n];
    }
    float ZB[3][3][3], ZD[3][3][3];
#pragma nounroll
    for (int row = 0; row < 9; row++) {
      const int qb = row % 3, qc = row / 3;
      const Coef1 cb = coef1(qb), cc = coef1(qc);
      __builtin_amdgcn_s_waitcnt(WAIT_VM0);
      if (qb == 0) slab_z32(Us, base, cc, ZB, ZD);
      float ur0[3][3], ur1[3][3], ur2[3][3];
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
          ur0[c][i] = fmac(cb.b[2], ZB[c][i][2], fmac(cb.b[1], ZB[c][i][1], static_cast<float>(cb.b[0]) * ZB[c][i][0]));
          ur1[c][i] = fmac(cb.d[2], ZB[c][i][2], fmac(cb.d[1], ZB[c][i][1], static_cast<float>(cb.d[0]) * ZB[c][i][0]));
          ur2[c][i] = fmac(cb.b[2], ZD[c][i][2], fmac(cb.b[1], ZD[c][i][1], static_cast<float>(cb.b[0]) * ZD[c][i][0]));
        }
      float P0[3][3], P1[3][3], P2[3][3];
      zero33(P0); zero33(P1); zero33(P2);
      // each point's factors come from LDS just before use; the next row's copy starts once
      // the three points have read theirs
#pragma unroll
      for (int qa = 0; qa < 3; qa++) {
        float F9[9], B4[4];
#pragma unroll
        for (int n = 0; n < 9; n++) F9[n] = Fs[(9 * qa + n) * BEL + lane];
#pragma unroll
        for (int n = 0; n < 4; n++) B4[n] = Bs[(4 * qa + n) * BEL + lane];
        if (qa == 0) stokes_qp32<0>(ur0, ur1, ur2, F9, B4, pe, P0, P1, P2, Yp);
        if (qa == 1) stokes_qp32<1>(ur0, ur1, ur2, F9, B4, pe, P0, P1, P2, Yp);
        if (qa == 2) stokes_qp32<2>(ur0, ur1, ur2, F9, B4, pe, P0, P1, P2, Yp);
      }
      if (row < 8) {
        __builtin_amdgcn_s_waitcnt(WAIT_LGKM0);
        lds_copy_async<FROW * 4>(fsrc + (row + 1) * FROW, Fs, lane);
        lds_copy_async<BROW * 4>(bsrc + (row + 1) * BROW, Bs, lane);
      }
      row_fence();
      acc_row_elem_rt(cb, cc, P0, P1, P2, Y);
      row_fence();
    }
    if (active) {
      float *o = yp + 4 * e;
#pragma unroll
      for (int i = 0; i < 4; i++) o[i] = add_mode ? o[i] + Yp[i] : Yp[i];
    }
  }
  __syncthreads();
  zero_tile(Ts, lane);
  __syncthreads();
  if (active) emit_element(Ts, base, Y);
  __syncthreads();
  if (color >= 0)
    store_tile(b, br, t, Ts, yu, add_mode != 0, lane);
  else
    store_tile_own(b, br, t, Ts, yu, facebuf, add_mode != 0, lane);
}

// FP64 geometry -> FP32 factors, with validation of det > 0 and eta > 0.
//   fac  [brick][q][9][lane]: F = sqrt(w eta / det) adj(J)
//   facb [brick][q][4][lane] (optional, for the FP32 full block): sqrt(w det / eta) (so that
//        w adj(J) = facb0 * F) and the P1GRel pressure basis values N1..N3 at the point.
template <int qa, int qb, int qc>
__device__ __forceinline__ void fac_qp(const double xr0[3][3], const double xr1[3][3], const double xr2[3][3],
                                       const double *eta27, const PBasis *pb, float *__restrict__ fp,
                                       float *__restrict__ fbp, int &bad) {
  double J[3][3], A[3][3];
  qp_grad<qa>(xr0, xr1, xr2, J);
  double det = adjugate(J, A);
  constexpr int q = qa + 3 * qb + 9 * qc;
  constexpr double w = gw1(qa) * gw1(qb) * gw1(qc);
  double et = eta27[q];
  if (!(det > 0.0) || !isfinite(det)) bad |= 1;
  if (!(et > 0.0) || !isfinite(et)) bad |= 2;
  double s = (bad == 0) ? sqrt(w * et / det) : 0.0;
#pragma unroll
  for (int m = 0; m < 9; m++) fp[(q * 9 + m) * BEL] = (float)(s * A[m / 3][m % 3]);
  if (fbp) {
    double x[3];
    qp_value<qa>(xr0, x);
    fbp[(q * 4) * BEL] = (float)((bad == 0) ? sqrt(w * det / et) : 0.0);
#pragma unroll
    for (int d = 0; d < 3; d++) fbp[(q * 4 + 1 + d) * BEL] = (float)((x[d] - pb->avg[d]) * pb->inv[d]);
  }
}

template <int qb, int qc>
__device__ __forceinline__ void fac_row(const double *Xs, int base, const double *eta27, const PBasis *pb, float *fp,
                                        float *fbp, int &bad) {
  double xr0[3][3], xr1[3][3], xr2[3][3];
  row_data<qb, qc>(Xs, base, xr0, xr1, xr2);
  fac_qp<0, qb, qc>(xr0, xr1, xr2, eta27, pb, fp, fbp, bad);
  fac_qp<1, qb, qc>(xr0, xr1, xr2, eta27, pb, fp, fbp, bad);
  fac_qp<2, qb, qc>(xr0, xr1, xr2, eta27, pb, fp, fbp, bad);
  row_fence();
}

__global__ __launch_bounds__(BEL) void k_build_fac32(Box b, const double *__restrict__ coords,
                                                     const double *__restrict__ eta_bm, float *__restrict__ fac,
                                                     float *__restrict__ facb, int *__restrict__ err) {
  __shared__ double Xs[3 * TCS];
  const int lane = threadIdx.x;
  const BrickId br = brick_lin(b, blockIdx.x);
  const Tile t = tile_of(b, br);
  load_tile(b, t, br, coords, (const uint8_t *)nullptr, Xs, lane);
  __syncthreads();
  if (!lane_active(t, lane)) return;
  const int base = lane_base(lane);
  double eta27[27];
  const double *ep = eta_bm + (int64_t)br.lin * 27 * BEL + lane;
#pragma unroll
  for (int q = 0; q < 27; q++) eta27[q] = ep[q * BEL];
  float *fp = fac + (int64_t)br.lin * 243 * BEL + lane;
  float *fbp = facb ? facb + (int64_t)br.lin * 108 * BEL + lane : nullptr;
  PBasis pbv;
  if (facb) pbv = pbasis_setup(Xs, base);
  const PBasis *pb = &pbv;
  int bad = 0;
  fac_row<0, 0>(Xs, base, eta27, pb, fp, fbp, bad); fac_row<1, 0>(Xs, base, eta27, pb, fp, fbp, bad);
  fac_row<2, 0>(Xs, base, eta27, pb, fp, fbp, bad); fac_row<0, 1>(Xs, base, eta27, pb, fp, fbp, bad);
  fac_row<1, 1>(Xs, base, eta27, pb, fp, fbp, bad); fac_row<2, 1>(Xs, base, eta27, pb, fp, fbp, bad);
  fac_row<0, 2>(Xs, base, eta27, pb, fp, fbp, bad); fac_row<1, 2>(Xs, base, eta27, pb, fp, fbp, bad);
  fac_row<2, 2>(Xs, base, eta27, pb, fp, fbp, bad);
  if (bad & 1) atomicOr(err, ERRBIT_MESH);
  if (bad & 2) atomicOr(err, ERRBIT_VISCOSITY);
}

// Natural eta[27e+q] -> brick-major [brick][q][lane]; validates eta > 0 and finite.
__global__ void k_eta_to_bm(Box b, const double *__restrict__ eta, double *__restrict__ eta_bm,
                            int *__restrict__ err) {
  const int lane = threadIdx.x & (BEL - 1);
  const int64_t blin = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / BEL;
  const int64_t nb = (int64_t)b.nbx * b.nby * b.nbz;
  if (blin >= nb) return;
  BrickId br = brick_lin(b, (int)blin);
  Tile t = tile_of(b, br);
  double *dst = eta_bm + blin * 27 * BEL + lane;
  if (!lane_active(t, lane)) {
    for (int q = 0; q < 27; q++) dst[q * BEL] = 1.0;
    return;
  }
  int64_t e = lane_element(b, t, lane);
  int bad = 0;
  for (int q = 0; q < 27; q++) {
    double v = eta[27 * e + q];
    if (!(v > 0.0) || !isfinite(v)) bad = 1;
    dst[q * BEL] = v;
  }
  if (bad) atomicOr(err, ERRBIT_VISCOSITY);
}

// Mesh validation: det J > 0 and finite at every quadrature point.
template <int qb, int qc>
__device__ __forceinline__ void check_row(const double *Xs, int base, int &bad) {
  double xr0[3][3], xr1[3][3], xr2[3][3], J[3][3], A[3][3], det;
  row_data<qb, qc>(Xs, base, xr0, xr1, xr2);
  qp_grad<0>(xr0, xr1, xr2, J); det = adjugate(J, A); if (!(det > 0.0) || !isfinite(det)) bad = 1;
  qp_grad<1>(xr0, xr1, xr2, J); det = adjugate(J, A); if (!(det > 0.0) || !isfinite(det)) bad = 1;
  qp_grad<2>(xr0, xr1, xr2, J); det = adjugate(J, A); if (!(det > 0.0) || !isfinite(det)) bad = 1;
  row_fence();
}

__global__ __launch_bounds__(BEL) void k_check_mesh(Box b, const double *__restrict__ coords, int *__restrict__ err) {
  __shared__ double Xs[3 * TCS];
  const int lane = threadIdx.x;
  const BrickId br = brick_lin(b, blockIdx.x);
  const Tile t = tile_of(b, br);
  load_tile(b, t, br, coords, (const uint8_t *)nullptr, Xs, lane);
  __syncthreads();
  if (!lane_active(t, lane)) return;
  const int base = lane_base(lane);
  int bad = 0;
  check_row<0, 0>(Xs, base, bad); check_row<1, 0>(Xs, base, bad); check_row<2, 0>(Xs, base, bad);
  check_row<0, 1>(Xs, base, bad); check_row<1, 1>(Xs, base, bad); check_row<2, 1>(Xs, base, bad);
  check_row<0, 2>(Xs, base, bad); check_row<1, 2>(Xs, base, bad); check_row<2, 2>(Xs, base, bad);
  if (bad) atomicOr(err, ERRBIT_MESH);
}

// ------------------------------------------------------------------ A21
// y_p[4e+i] += -sum_q w det N^p_i (div u)_q ; div u = trace(M)/det.

template <int qa, int qb, int qc>
__device__ __forceinline__ void a21_qp(const double xr0[3][3], const double xr1[3][3], const double xr2[3][3],
                                       const double ur0[3][3], const double ur1[3][3], const double ur2[3][3],
                                       const PBasis &pb, double Yp[4]) {
  double J[3][3], A[3][3], g[3][3], x[3];
  qp_grad<qa>(xr0, xr1, xr2, J);
  adjugate(J, A);
  qp_grad<qa>(ur0, ur1, ur2, g);
  qp_value<qa>(xr0, x);
  // trace(M) = sum_k sum_m g[m][k] A[k][m] = det * div u
  double tr = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) tr += g[0][k] * A[k][0] + g[1][k] * A[k][1] + g[2][k] * A[k][2];
  constexpr double w = gw1(qa) * gw1(qb) * gw1(qc);
  double f = -w * tr;
  Yp[0] += f;
#pragma unroll
  for (int d = 0; d < 3; d++) Yp[1 + d] += f * ((x[d] - pb.avg[d]) * pb.inv[d]);
}

template <int qb, int qc>
__device__ __forceinline__ void a21_row(const double *Xs, const double *Us, int base, const PBasis &pb, double Yp[4]) {
  double xr0[3][3], xr1[3][3], xr2[3][3], ur0[3][3], ur1[3][3], ur2[3][3];
  row_data<qb, qc>(Xs, base, xr0, xr1, xr2);
  row_data<qb, qc>(Us, base, ur0, ur1, ur2);
  a21_qp<0, qb, qc>(xr0, xr1, xr2, ur0, ur1, ur2, pb, Yp);
  a21_qp<1, qb, qc>(xr0, xr1, xr2, ur0, ur1, ur2, pb, Yp);
  a21_qp<2, qb, qc>(xr0, xr1, xr2, ur0, ur1, ur2, pb, Yp);
  row_fence();
}

__global__ __launch_bounds__(BEL) void k_a21(Box b, const double *__restrict__ coords, const double *__restrict__ u,
                                             const uint8_t *__restrict__ mask, double *__restrict__ yp, int add_mode) {
  __shared__ double Xs[3 * TCS];
  __shared__ double Us[3 * TCS];
  const int lane = threadIdx.x;
  const BrickId br = brick_lin(b, blockIdx.x);
  const Tile t = tile_of(b, br);
  load_tile2(b, t, br, coords, Xs, u, mask, Us, lane);
  __syncthreads();
  if (!lane_active(t, lane)) return;
  const int base = lane_base(lane);
  const PBasis pb = pbasis_setup(Xs, base);
  double Yp[4] = {0.0, 0.0, 0.0, 0.0};
  a21_row<0, 0>(Xs, Us, base, pb, Yp); a21_row<1, 0>(Xs, Us, base, pb, Yp); a21_row<2, 0>(Xs, Us, base, pb, Yp);
  a21_row<0, 1>(Xs, Us, base, pb, Yp); a21_row<1, 1>(Xs, Us, base, pb, Yp); a21_row<2, 1>(Xs, Us, base, pb, Yp);
  a21_row<0, 2>(Xs, Us, base, pb, Yp); a21_row<1, 2>(Xs, Us, base, pb, Yp); a21_row<2, 2>(Xs, Us, base, pb, Yp);
  const int64_t e = lane_element(b, t, lane);
  double *o = yp + 4 * e;
#pragma unroll
  for (int i = 0; i < 4; i++) o[i] = add_mode ? o[i] + Yp[i] : Yp[i];
}

// ------------------------------------------------------------------ A12
// y_u[n,l] += -sum_q w p_q (A^T)... : dv[k][l] = -w p A[l][k] (QuadratureAction_A12_AVX).

template <int qa, int qb, int qc>
__device__ __forceinline__ void a12_qp(const double xr0[3][3], const double xr1[3][3], const double xr2[3][3],
                                       const PBasis &pb, const double pe[4], double RB[3][3][3], double RD[3][3][3]) {
  double J[3][3], A[3][3], x[3], dv[3][3];
  qp_grad<qa>(xr0, xr1, xr2, J);
  adjugate(J, A);
  qp_value<qa>(xr0, x);
  double pq = pe[0];
#pragma unroll
  for (int d = 0; d < 3; d++) pq += pe[1 + d] * ((x[d] - pb.avg[d]) * pb.inv[d]);
  constexpr double w = gw1(qa) * gw1(qb) * gw1(qc);
  double f = -w * pq;
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int l = 0; l < 3; l++) dv[k][l] = f * A[l][k];
  acc_direct<qa, qb>(dv, RB, RD);
}

template <int qb, int qc>
__device__ __forceinline__ void a12_row(const double *Xs, int base, const PBasis &pb, const double pe[4],
                                        double RB[3][3][3], double RD[3][3][3]) {
  double xr0[3][3], xr1[3][3], xr2[3][3];
  row_data<qb, qc>(Xs, base, xr0, xr1, xr2);
  a12_qp<0, qb, qc>(xr0, xr1, xr2, pb, pe, RB, RD);
  a12_qp<1, qb, qc>(xr0, xr1, xr2, pb, pe, RB, RD);
  a12_qp<2, qb, qc>(xr0, xr1, xr2, pb, pe, RB, RD);
  row_fence();
}

template <int qc>
__device__ __forceinline__ void a12_slab(const double *Xs, int base, const PBasis &pb, const double pe[4],
                                         double Y[3][3][3][3]) {
  double RB[3][3][3], RD[3][3][3];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) RB[c][j][i] = RD[c][j][i] = 0.0;
  a12_row<0, qc>(Xs, base, pb, pe, RB, RD);
  a12_row<1, qc>(Xs, base, pb, pe, RB, RD);
  a12_row<2, qc>(Xs, base, pb, pe, RB, RD);
  acc_elem<qc>(RB, RD, Y);
  row_fence();
}

__global__ __launch_bounds__(BEL) void k_a12(Box b, int color, const double *__restrict__ coords,
                                             const double *__restrict__ p, double *__restrict__ y, int add_mode) {
  __shared__ double Xs[3 * TCS];
  const int lane = threadIdx.x;
  const BrickId br = brick_of(b, color, blockIdx.x);
  const Tile t = tile_of(b, br);
  load_tile(b, t, br, coords, (const uint8_t *)nullptr, Xs, lane);
  __syncthreads();
  const int base = lane_base(lane);
  const bool active = lane_active(t, lane);
  double Y[3][3][3][3];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) Y[c][k][j][i] = 0.0;
  if (active) {
    const PBasis pb = pbasis_setup(Xs, base);
    const int64_t e = lane_element(b, t, lane);
    double pe[4] = {p[4 * e], p[4 * e + 1], p[4 * e + 2], p[4 * e + 3]};
    a12_slab<0>(Xs, base, pb, pe, Y);
    a12_slab<1>(Xs, base, pb, pe, Y);
    a12_slab<2>(Xs, base, pb, pe, Y);
  }
  __syncthreads();
  zero_tile(Xs, lane);
  __syncthreads();
  if (active) emit_element(Xs, base, Y);
  __syncthreads();
  store_tile(b, br, t, Xs, y, add_mode != 0, lane);
}

// ---------------------------------------------------------------- body force
// y[3n+d] += sum_e sum_q w_q det J_q N_n(x_q) fu[e][q][d]: the Be term of FormFunctionLocal_U
// (src/stokes_form_function.c:219-224), the momentum load of buoyancy problems.  One launch over
// all bricks with the owner-writes store (face slots, then k_gather_faces).  Per element the 27
// weights w det J come from the coordinate tile; its LDS then becomes the output tile.  Then, one
// z-slab of quadrature points (qc) at a time, each lane reads its element's 27 values of the slab
// (consecutive in fu, 16-byte loads), scales them by w det, applies the transposed Q2
// interpolation along x and y and adds the slab's z factor times that into the output tile.
// With nothing but the weights and two slabs held, two bricks run per SIMD, so one brick's loads
// overlap the other's work.

// the lane's element's 27 values of slab qc (zeros outside the box, never used): 13 16-byte loads
// and one 8-byte load, the single one first or last as the slab's address is or is not 16-byte
// aligned
__device__ __forceinline__ void bf_slab_load(const double *__restrict__ fe, bool active, int qc, double r[27]) {
  const double *p = fe + 27 * qc;
  const bool al = ((uintptr_t)p & 15) == 0;
  const double2 *__restrict__ p4 = (const double2 *)(al ? p : p + 1);
  const double *__restrict__ p1 = al ? p + 26 : p;
  double2 v[13];
  double w = 0.0;
#pragma unroll
  for (int k = 0; k < 13; k++) v[k] = active ? p4[k] : make_double2(0.0, 0.0);
  if (active) w = *p1;
#pragma unroll
  for (int k = 0; k < 27; k++) {
    const double a = k < 26 ? (k & 1 ? v[k >> 1].y : v[k >> 1].x) : w;
    const double m = k == 0 ? w : ((k - 1) & 1 ? v[(k - 1) >> 1].y : v[(k - 1) >> 1].x);
    r[k] = al ? a : m;
  }
}

// geometry weight w det J (times sc) of quadrature point (qa, qb, qc)
template <int qa, int qb, int qc>
__device__ __forceinline__ void bf_weight(const double xr0[3][3], const double xr1[3][3], const double xr2[3][3],
                                          double sc, double wd[27]) {
  double J[3][3], A[3][3];
  qp_grad<qa>(xr0, xr1, xr2, J);
  const double det = adjugate(J, A);
  constexpr double w = gw1(qa) * gw1(qb) * gw1(qc);
  wd[qa + 3 * qb + 9 * qc] = sc * (w * det);
}

template <int qb, int qc>
__device__ __forceinline__ void bf_row(const double *Xs, int base, double sc, double wd[27]) {
  double xr0[3][3], xr1[3][3], xr2[3][3];
  row_data<qb, qc>(Xs, base, xr0, xr1, xr2);
  bf_weight<0, qb, qc>(xr0, xr1, xr2, sc, wd);
  bf_weight<1, qb, qc>(xr0, xr1, xr2, sc, wd);
  bf_weight<2, qb, qc>(xr0, xr1, xr2, sc, wd);
}

// slab qc into the output tile: F = w det fu at the slab's 9 points, B^T along x, then y, then
// B(qc, k) into node layer k (ds_add_f64: the 64 addresses of one instruction are distinct, so
// the order of the sums is the program order)
template <int qc>
__device__ __forceinline__ void bf_slab(double *Ys, int base, const double r[27], const double wd[27]) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double T1[3][3];  // [qb][i]
#pragma unroll
    for (int qb = 0; qb < 3; qb++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        double a = 0.0;
        bool first = true;
#pragma unroll
        for (int qa = 0; qa < 3; qa++) {
          if (!nzB(qa, i)) continue;
          const double f = wd[qa + 3 * qb + 9 * qc] * r[3 * (qa + 3 * qb) + d];
          a = first ? gB(qa, i) * f : fma(gB(qa, i), f, a);
          first = false;
        }
        T1[qb][i] = a;
      }
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        double a = 0.0;
        bool first = true;
#pragma unroll
        for (int qb = 0; qb < 3; qb++) {
          if (!nzB(qb, j)) continue;
          a = first ? gB(qb, j) * T1[qb][i] : fma(gB(qb, j), T1[qb][i], a);
          first = false;
        }
#pragma unroll
        for (int k = 0; k < 3; k++)
          if (nzB(qc, k)) lds_add(&Ys[base + toff(d, i, j, k)], gB(qc, k) * a);
      }
  }
}

__global__ __launch_bounds__(BEL) void k_body_force(Box b, const double *__restrict__ coords,
                                                    const double *__restrict__ fu, double *__restrict__ y,
                                                    int add_mode, double sc, double *__restrict__ facebuf) {
  __shared__ double Xs[3 * TCS];
  const int lane = threadIdx.x;
  const BrickId br = brick_xcd(b, blockIdx.x);
  const Tile t = tile_of(b, br);
  const bool active = lane_active(t, lane);
  const double *fe = fu + 81 * (active ? lane_element(b, t, lane) : 0);
  double r0[27], r1[27];
  bf_slab_load(fe, active, 0, r0);
  load_tile(b, t, br, coords, (const uint8_t *)nullptr, Xs, lane);
  __syncthreads();
  const int base = lane_base(lane);
  double wd[27];
  if (active) {
    bf_row<0, 0>(Xs, base, sc, wd); row_fence(); bf_row<1, 0>(Xs, base, sc, wd); row_fence();
    bf_row<2, 0>(Xs, base, sc, wd); row_fence(); bf_row<0, 1>(Xs, base, sc, wd); row_fence();
    bf_row<1, 1>(Xs, base, sc, wd); row_fence(); bf_row<2, 1>(Xs, base, sc, wd); row_fence();
    bf_row<0, 2>(Xs, base, sc, wd); row_fence(); bf_row<1, 2>(Xs, base, sc, wd); row_fence();
    bf_row<2, 2>(Xs, base, sc, wd);
  }
  bf_slab_load(fe, active, 1, r1);
  __syncthreads();  // the coordinate tile is read: its LDS becomes the output tile
  zero_tile(Xs, lane);
  __syncthreads();
  if (active) bf_slab<0>(Xs, base, r0, wd);
  bf_slab_load(fe, active, 2, r0);
  if (active) bf_slab<1>(Xs, base, r1, wd);
  if (active) bf_slab<2>(Xs, base, r0, wd);
  __syncthreads();
  store_tile_own(b, br, t, Xs, y, facebuf, add_mode != 0, lane);
}

// ---------------------------------------------------------------- diagonal
// diag(n,x) += sum_q w det eta (2 Nx^2 + Ny^2 + Nz^2) (cyclic for y, z), the form of
// MatMultMF_Stokes_MixedFEM3d_diagB11 (src/stokes_q2p1_mf_operators_diag_def.c:120-205).
// With h = A grad_xi N = det grad_x N: w det eta Nx^2 = (w eta / det) h_x^2.

template <int qa, int qb, int qc>
__device__ __forceinline__ void diag_qp(const double xr0[3][3], const double xr1[3][3], const double xr2[3][3],
                                        const double *eta9, double Y[3][3][3][3]) {
  double J[3][3], A[3][3];
  qp_grad<qa>(xr0, xr1, xr2, J);
  double det = adjugate(J, A);
  constexpr double w = gw1(qa) * gw1(qb) * gw1(qc);
  double s = (w * eta9[qa + 3 * qb]) * rcp64(det);
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const double n0 = gD(qa, i) * gB(qb, j) * gB(qc, k);
        const double n1 = gB(qa, i) * gD(qb, j) * gB(qc, k);
        const double n2 = gB(qa, i) * gB(qb, j) * gD(qc, k);
        if (n0 != 0.0 || n1 != 0.0 || n2 != 0.0) {
          double hx = A[0][0] * n0 + A[0][1] * n1 + A[0][2] * n2;
          double hy = A[1][0] * n0 + A[1][1] * n1 + A[1][2] * n2;
          double hz = A[2][0] * n0 + A[2][1] * n1 + A[2][2] * n2;
          double ax = hx * hx, ay = hy * hy, az = hz * hz, sum = ax + ay + az;
          Y[0][k][j][i] += s * (sum + ax);
          Y[1][k][j][i] += s * (sum + ay);
          Y[2][k][j][i] += s * (sum + az);
        }
      }
}

template <int qb, int qc>
__device__ __forceinline__ void diag_row(const double *Xs, int base, const double *eta9, double Y[3][3][3][3]) {
  double xr0[3][3], xr1[3][3], xr2[3][3];
  row_data<qb, qc>(Xs, base, xr0, xr1, xr2);
  diag_qp<0, qb, qc>(xr0, xr1, xr2, eta9, Y);
  row_fence();
  diag_qp<1, qb, qc>(xr0, xr1, xr2, eta9, Y);
  row_fence();
  diag_qp<2, qb, qc>(xr0, xr1, xr2, eta9, Y);
  row_fence();
}

__global__ __launch_bounds__(BEL) void k_diag(Box b, int color, const double *__restrict__ coords,
                                              const double *__restrict__ eta_bm, double *__restrict__ d, int add_mode) {
  __shared__ double Xs[3 * TCS];
  const int lane = threadIdx.x;
  const BrickId br = brick_of(b, color, blockIdx.x);
  const Tile t = tile_of(b, br);
  load_tile(b, t, br, coords, (const uint8_t *)nullptr, Xs, lane);
  __syncthreads();
  const int base = lane_base(lane);
  const bool active = lane_active(t, lane);
  double Y[3][3][3][3];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) Y[c][k][j][i] = 0.0;
  if (active) {
    const double *ep = eta_bm + (int64_t)br.lin * 27 * BEL + lane;
    double e9[9];
    load9(ep, 0, e9);
    diag_row<0, 0>(Xs, base, e9, Y); diag_row<1, 0>(Xs, base, e9, Y); diag_row<2, 0>(Xs, base, e9, Y);
    load9(ep, 1, e9);
    diag_row<0, 1>(Xs, base, e9, Y); diag_row<1, 1>(Xs, base, e9, Y); diag_row<2, 1>(Xs, base, e9, Y);
    load9(ep, 2, e9);
    diag_row<0, 2>(Xs, base, e9, Y); diag_row<1, 2>(Xs, base, e9, Y); diag_row<2, 2>(Xs, base, e9, Y);
  }
  __syncthreads();
  zero_tile(Xs, lane);
  __syncthreads();
  if (active) emit_element(Xs, base, Y);
  __syncthreads();
  store_tile(b, br, t, Xs, d, add_mode != 0, lane);
}

// ------------------------------------------------------------- BC helpers
// Dirichlet rows: y[idx] = x[idx] (mode 0, BCListInsertDirichlet_MatMult), 0 (mode 1,
// BCListInsertZero), 1 (mode 2, BCListInsertValueIntoDirichletSlot) or x[idx] - v[idx] with v
// null meaning 0 (mode 3, BCListResidualDirichlet).
template <typename T>
__global__ void k_bc_rows(int64_t n, const int64_t *__restrict__ idx, const T *__restrict__ x, T *__restrict__ y,
                          int mode, const T *__restrict__ v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t k = idx[i];
  y[k] = mode == 0 ? x[k] : (mode == 1 ? T(0) : (mode == 2 ? T(1) : x[k] - (v ? v[k] : T(0))));
}

// Non-finite scan of an input vector (sets ERRBIT_NONFINITE).
template <typename T>
__global__ void k_scan_finite(int64_t n, const T *__restrict__ x, int *__restrict__ err) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int bad = 0;
  for (; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (!isfinite(x[i])) bad = 1;
  if (bad) atomicOr(err, ERRBIT_NONFINITE);
}

// ------------------------------------------------------------ launchers
static inline int nbricks_color(const Box &b, int color) {
  int cx = color & 1, cy = (color >> 1) & 1, cz = (color >> 2) & 1;
  return ((b.nbx - cx + 1) >> 1) * ((b.nby - cy + 1) >> 1) * ((b.nbz - cz + 1) >> 1);
}
static inline int nbricks(const Box &b) { return b.nbx * b.nby * b.nbz; }

// Second pass of the uncoloured output path: every brick adds into the nodes of its lower
// faces the face slots of its up to 7 lower neighbours (store_tile_own), in increasing brick
// order after the owner's own value.  The neighbour planes, edges and corner are staged in
// LDS by one batch of independent loads (the XCD-aware brick order keeps most of them on the
// same L2 as their writer); then each thread takes up to 3 of the brick's at most 3 x 3 x TN x TN
// face values.
template <typename T>
__device__ __forceinline__ void gather_faces_brick(const Box &b, const BrickId &br, const T *__restrict__ face,
                                                   T *__restrict__ y) {
  constexpr int F2 = TN * TN, UP = 2 * BE;
  constexpr int OFX = 0, OFY = 3 * F2, OFZ = 6 * F2, OXY = 9 * F2, OXZ = OXY + 3 * TN, OYZ = OXZ + 3 * TN,
                OC = OYZ + 3 * TN, NS = OC + 3;
  __shared__ T sh[NS];
  const bool hx = br.bx > 0, hy = br.by > 0, hz = br.bz > 0;
  if (!(hx || hy || hz)) return;
  const Tile t = tile_of(b, br);
  const int tid = threadIdx.x;
  const int64_t fx = FACEVALS, fy = (int64_t)b.nbx * FACEVALS, fz = (int64_t)b.nbx * b.nby * FACEVALS;
  const T *fb = face + (int64_t)br.lin * FACEVALS;  // own slot: the fallback address of absent neighbours
  // slot value (plane, row a, column b, component c) sits at 3 * (plane * F2 + a * TN + b) + c
  const int g = min(tid, 3 * F2 - 1), gc = g / F2, gr = g - gc * F2;
  const T vx = fb[(hx ? -fx : 0) + 3 * gr + gc];           // X plane of bx-1: (UP, y, z), gr = z*TN + y
  const T vy = fb[(hy ? -fy : 0) + 3 * (F2 + gr) + gc];    // Y plane of by-1: (x, UP, z), gr = z*TN + x
  const T vz = fb[(hz ? -fz : 0) + 3 * (2 * F2 + gr) + gc];  // Z plane of bz-1: (x, y, UP), gr = y*TN + x
  const int e = min(tid, 3 * TN - 1), ec = e / TN, ep = e - ec * TN;
  const T exy = fb[(hx && hy ? -fx - fy : 0) + 3 * (ep * TN + UP) + ec];          // X plane: (UP, UP, z = ep)
  const T exz = fb[(hx && hz ? -fx - fz : 0) + 3 * (UP * TN + ep) + ec];          // X plane: (UP, y = ep, UP)
  const T eyz = fb[(hy && hz ? -fy - fz : 0) + 3 * (F2 + UP * TN + ep) + ec];     // Y plane: (x = ep, UP, UP)
  const T cxyz = fb[(hx && hy && hz ? -fx - fy - fz : 0) + 3 * (UP * TN + UP) + min(tid, 2)];  // X plane corner
  if (tid < 3 * F2) {
    sh[OFX + tid] = vx;
    sh[OFY + tid] = vy;
    sh[OFZ + tid] = vz;
  }
  if (tid < 3 * TN) {
    sh[OXY + tid] = exy;
    sh[OXZ + tid] = exz;
    sh[OYZ + tid] = eyz;
  }
  if (tid < 3) sh[OC + tid] = cxyz;
  __syncthreads();
  // the brick's nodes on its lower faces: plane x = 0, then y = 0 (x >= x0), then z = 0
  // (x >= x0, y >= y0), each node once
  const int ox = br.bx == b.nbx - 1 ? t.tnx : UP;
  const int oy = br.by == b.nby - 1 ? t.tny : UP;
  const int oz = br.bz == b.nbz - 1 ? t.tnz : UP;
  const int x0 = hx ? 1 : 0, y0 = hy ? 1 : 0;
  const int nX = hx ? oy * oz : 0, nY = hy ? (ox - x0) * oz : 0, nZ = hz ? (ox - x0) * (oy - y0) : 0;
  const int nval = 3 * (nX + nY + nZ);
  int64_t gi[3];
  int ix[3], iy[3], iz[3], cc[3];
  T own[3];
#pragma unroll
  for (int m = 0; m < 3; m++) {
    const int v = min(tid + 256 * m, nval - 1);
    const int n = v / 3;
    cc[m] = v - 3 * n;
    if (n < nX) {
      ix[m] = 0; iy[m] = n % oy; iz[m] = n / oy;
    } else if (n < nX + nY) {
      const int q = n - nX;
      ix[m] = x0 + q % (ox - x0); iy[m] = 0; iz[m] = q / (ox - x0);
    } else {
      const int q = n - nX - nY;
      ix[m] = x0 + q % (ox - x0); iy[m] = y0 + q / (ox - x0); iz[m] = 0;
    }
    gi[m] = 3 * (t.node0 + ix[m] + (int64_t)b.NX * (iy[m] + (int64_t)b.NY * iz[m])) + cc[m];
    own[m] = y[gi[m]];
  }
#pragma unroll
  for (int m = 0; m < 3; m++) {
    if (tid + 256 * m >= nval) break;
    const int c = cc[m], X = ix[m], Y = iy[m], Z = iz[m];
    const bool nx = hx && X == 0, ny = hy && Y == 0, nz = hz && Z == 0;
    T acc = own[m];
    if (nz && ny && nx) acc += sh[OC + c];
    if (nz && ny) acc += sh[OYZ + c * TN + X];
    if (nz && nx) acc += sh[OXZ + c * TN + Y];
    if (nz) acc += sh[OFZ + c * F2 + Y * TN + X];
    if (ny && nx) acc += sh[OXY + c * TN + Z];
    if (ny) acc += sh[OFY + c * F2 + Z * TN + X];
    if (nx) acc += sh[OFX + c * F2 + Z * TN + Y];
    y[gi[m]] = acc;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather_faces(Box b, const T *__restrict__ face, T *__restrict__ y) {
  gather_faces_brick(b, brick_xcd(b, blockIdx.x), face, y);
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather_faces_batch(BatchArgs<T> a) {
  const int p = batch_problem(a.start, a.n, blockIdx.x);
  const Box b = a.box[p];
  const int local = (int)blockIdx.x - a.start[p];
  if (local >= b.nbx * b.nby * b.nbz) return;
  gather_faces_brick(b, brick_xcd(b, local), (const T *)a.face[p], a.y[p]);
}

template <typename T>
hipError_t launch_gather(const Box &b, const T *face, T *y, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_faces<T>, dim3(nbricks(b)), dim3(256), 0, s, b, face, y);
  return hipGetLastError();
}

hipError_t launch_a11_fp64(const Box &b, const double *coords, const double *u, const uint8_t *mask,
                           const double *eta_bm, double *y, bool add, hipStream_t s, double *facebuf, bool gather,
                           const ChebOut<double> *co) {
  if (facebuf) {
    if (co) {
      hipLaunchKernelGGL((k_a11_fp64_t<false, true>), dim3(nbricks(b)), dim3(BEL), 0, s, b, -1, coords, u, mask, eta_bm,
                         y, 0, facebuf, (const double *)nullptr, (double *)nullptr, (const double *)nullptr, *co);
      return hipGetLastError();
    }
    hipLaunchKernelGGL((k_a11_fp64_t<false, false>), dim3(nbricks(b)), dim3(BEL), 0, s, b, -1, coords, u, mask, eta_bm, y,
                       add ? 1 : 0, facebuf, (const double *)nullptr, (double *)nullptr, (const double *)nullptr,
                       ChebOut<double>{});
    return gather ? launch_gather(b, (const double *)facebuf, y, s) : hipGetLastError();
  }
  for (int c = 0; c < 8; c++) {
    int nb = nbricks_color(b, c);
    if (nb == 0) continue;
    hipLaunchKernelGGL((k_a11_fp64_t<false, false>), dim3(nb), dim3(BEL), 0, s, b, c, coords, u, mask, eta_bm, y, add ? 1 : 0,
                       (double *)nullptr, (const double *)nullptr, (double *)nullptr, (const double *)nullptr,
                       ChebOut<double>{});
  }
  return hipGetLastError();
}

hipError_t launch_gather_faces(const Box &b, const float *face, float *y, hipStream_t s) {
  return launch_gather(b, face, y, s);
}
hipError_t launch_gather_faces(const Box &b, const double *face, double *y, hipStream_t s) {
  return launch_gather(b, face, y, s);
}

hipError_t launch_a11_fp64_split(const Box &b, const double *coords, const double *u, const uint8_t *mask,
                                 const double *eta_bm, double *y, bool add, hipStream_t s, double *facebuf,
                                 bool gather) {
  if (!facebuf) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_a11_fp64_split, dim3(nbricks(b)), dim3(3 * BEL), 0, s, b, coords, u, mask, eta_bm, y,
                     add ? 1 : 0, facebuf);
  return gather ? launch_gather(b, (const double *)facebuf, y, s) : hipGetLastError();
}

hipError_t launch_a11_fp32_split(const Box &b, const float *u, const uint8_t *mask, const float *fac, float *y,
                                 bool add, hipStream_t s, float *facebuf, bool gather) {
  if (!facebuf) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_a11_fp32_split, dim3(nbricks(b)), dim3(3 * BEL), 0, s, b, u, mask, fac, y, add ? 1 : 0,
                     facebuf);
  return gather ? launch_gather(b, (const float *)facebuf, y, s) : hipGetLastError();
}

template <typename T>
hipError_t launch_a11_batch(const BatchArgs<T> &a, hipStream_t s) {
  const int nb = a.start[a.n];
  if constexpr (sizeof(T) == 8)
    hipLaunchKernelGGL(k_a11_fp64_batch, dim3(nb), dim3(BEL), 0, s, a);
  else
    hipLaunchKernelGGL(k_a11_fp32_batch, dim3(nb), dim3(BEL), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_gather_faces_batch<T>, dim3(nb), dim3(256), 0, s, a);
  return hipGetLastError();
}
template hipError_t launch_a11_batch<double>(const BatchArgs<double> &, hipStream_t);
template hipError_t launch_a11_batch<float>(const BatchArgs<float> &, hipStream_t);

hipError_t launch_stokes_fp64(const Box &b, const double *coords, const double *u, const double *p,
                              const uint8_t *mask, const double *eta_bm, double *yu, double *yp, const double *fp,
                              bool add, hipStream_t s, double *facebuf) {
  if (facebuf) {
    hipLaunchKernelGGL((k_a11_fp64_t<true, false>), dim3(nbricks(b)), dim3(BEL), 0, s, b, -1, coords, u, mask, eta_bm, yu,
                       add ? 1 : 0, facebuf, p, yp, fp, ChebOut<double>{});
    return launch_gather(b, (const double *)facebuf, yu, s);
  }
  for (int c = 0; c < 8; c++) {
    int nb = nbricks_color(b, c);
    if (nb == 0) continue;
    hipLaunchKernelGGL((k_a11_fp64_t<true, false>), dim3(nb), dim3(BEL), 0, s, b, c, coords, u, mask, eta_bm, yu,
                       add ? 1 : 0, (double *)nullptr, p, yp, fp, ChebOut<double>{});
  }
  return hipGetLastError();
}

hipError_t launch_a11_fp32(const Box &b, const float *u, const uint8_t *mask, const float *fac, float *y, bool add,
                           hipStream_t s, float *facebuf, bool gather, const ChebOut<float> *co) {
  if (facebuf) {
    if (co) {
      hipLaunchKernelGGL(k_a11_fp32<true>, dim3(nbricks(b)), dim3(BEL), 0, s, b, -1, u, mask, fac, y, 0, facebuf, *co);
      return hipGetLastError();
    }
    hipLaunchKernelGGL(k_a11_fp32<false>, dim3(nbricks(b)), dim3(BEL), 0, s, b, -1, u, mask, fac, y, add ? 1 : 0,
                       facebuf, ChebOut<float>{});
    return gather ? launch_gather(b, (const float *)facebuf, y, s) : hipGetLastError();
  }
  for (int c = 0; c < 8; c++) {
    int nb = nbricks_color(b, c);
    if (nb == 0) continue;
    hipLaunchKernelGGL(k_a11_fp32<false>, dim3(nb), dim3(BEL), 0, s, b, c, u, mask, fac, y, add ? 1 : 0,
                       (float *)nullptr, ChebOut<float>{});
  }
  return hipGetLastError();
}

hipError_t launch_build_fac32(const Box &b, const double *coords, const double *eta_bm, float *fac, float *facb,
                              int *err, hipStream_t s) {
  hipLaunchKernelGGL(k_build_fac32, dim3(nbricks(b)), dim3(BEL), 0, s, b, coords, eta_bm, fac, facb, err);
  return hipGetLastError();
}

hipError_t launch_a11_fp64g(const Box &b, const double *u, const uint8_t *mask, const double *geo, const double *eta_bm,
                            double *y, bool add, hipStream_t s, double *facebuf) {
  if (facebuf) {
    hipLaunchKernelGGL(k_a11_fp64g, dim3(nbricks(b)), dim3(BEL), 0, s, b, -1, u, mask, geo, eta_bm, y, add ? 1 : 0,
                       facebuf);
    return launch_gather(b, (const double *)facebuf, y, s);
  }
  for (int c = 0; c < 8; c++) {
    int nb = nbricks_color(b, c);
    if (nb == 0) continue;
    hipLaunchKernelGGL(k_a11_fp64g, dim3(nb), dim3(BEL), 0, s, b, c, u, mask, geo, eta_bm, y, add ? 1 : 0,
                       (double *)nullptr);
  }
  return hipGetLastError();
}

hipError_t launch_build_geo64(const Box &b, const double *coords, double *geo, hipStream_t s) {
  hipLaunchKernelGGL(k_build_geo64, dim3(nbricks(b)), dim3(BEL), 0, s, b, coords, geo);
  return hipGetLastError();
}

hipError_t launch_stokes_fp32(const Box &b, const float *u, const float *p, const uint8_t *mask, const float *fac,
                              const float *facb, float *yu, float *yp, bool add, hipStream_t s, float *facebuf) {
  if (facebuf) {
    hipLaunchKernelGGL(k_stokes_fp32, dim3(nbricks(b)), dim3(BEL), 0, s, b, -1, u, p, mask, fac, facb, yu, yp,
                       add ? 1 : 0, facebuf);
    return launch_gather(b, (const float *)facebuf, yu, s);
  }
  for (int c = 0; c < 8; c++) {
    int nb = nbricks_color(b, c);
    if (nb == 0) continue;
    hipLaunchKernelGGL(k_stokes_fp32, dim3(nb), dim3(BEL), 0, s, b, c, u, p, mask, fac, facb, yu, yp, add ? 1 : 0,
                       (float *)nullptr);
  }
  return hipGetLastError();
}

hipError_t launch_eta_to_bm(const Box &b, const double *eta, double *eta_bm, int *err, hipStream_t s) {
  int64_t threads = (int64_t)nbricks(b) * BEL;
  int blocks = (int)((threads + 255) / 256);
  hipLaunchKernelGGL(k_eta_to_bm, dim3(blocks), dim3(256), 0, s, b, eta, eta_bm, err);
  return hipGetLastError();
}

hipError_t launch_check_mesh(const Box &b, const double *coords, int *err, hipStream_t s) {
  hipLaunchKernelGGL(k_check_mesh, dim3(nbricks(b)), dim3(BEL), 0, s, b, coords, err);
  return hipGetLastError();
}

hipError_t launch_a21(const Box &b, const double *coords, const double *u, const uint8_t *mask, double *yp, bool add,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_a21, dim3(nbricks(b)), dim3(BEL), 0, s, b, coords, u, mask, yp, add ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_a12(const Box &b, const double *coords, const double *p, double *y, bool add, hipStream_t s) {
  for (int c = 0; c < 8; c++) {
    int nb = nbricks_color(b, c);
    if (nb == 0) continue;
    hipLaunchKernelGGL(k_a12, dim3(nb), dim3(BEL), 0, s, b, c, coords, p, y, add ? 1 : 0);
  }
  return hipGetLastError();
}

hipError_t launch_body_force(const Box &b, const double *coords, const double *fu, double *y, bool add,
                             double *facebuf, hipStream_t s, double scale) {
  hipLaunchKernelGGL(k_body_force, dim3(nbricks(b)), dim3(BEL), 0, s, b, coords, fu, y, add ? 1 : 0, scale, facebuf);
  hipError_t e = hipGetLastError();
  return e != hipSuccess ? e : launch_gather(b, (const double *)facebuf, y, s);
}

hipError_t launch_diag(const Box &b, const double *coords, const double *eta_bm, double *d, bool add, hipStream_t s) {
  for (int c = 0; c < 8; c++) {
    int nb = nbricks_color(b, c);
    if (nb == 0) continue;
    hipLaunchKernelGGL(k_diag, dim3(nb), dim3(BEL), 0, s, b, c, coords, eta_bm, d, add ? 1 : 0);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_bc_rows(int64_t n, const int64_t *idx, const T *x, T *y, int mode, hipStream_t s, const T *v) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bc_rows<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, idx, x, y, mode, v);
  return hipGetLastError();
}
template hipError_t launch_bc_rows<double>(int64_t, const int64_t *, const double *, double *, int, hipStream_t,
                                           const double *);
template hipError_t launch_bc_rows<float>(int64_t, const int64_t *, const float *, float *, int, hipStream_t,
                                          const float *);

template <typename T>
hipError_t launch_scan_finite(int64_t n, const T *x, int *err, hipStream_t s) {
  if (n == 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_scan_finite<T>, dim3((unsigned)blocks), dim3(256), 0, s, n, x, err);
  return hipGetLastError();
}
template hipError_t launch_scan_finite<double>(int64_t, const double *, int *, hipStream_t);
template hipError_t launch_scan_finite<float>(int64_t, const float *, int *, hipStream_t);

}  // namespace ptq2
