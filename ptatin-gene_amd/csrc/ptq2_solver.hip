// ptq2_solver.hip -- GMG-preconditioned FGMRES for the viscous block (include/ptq2_solver.h).
//
// Level l has its own ptq2_ctx (mesh, BC mask, viscosity).  Coarse meshes are the
// fine node lattice injected every second node (src/dmda_project_coords.c:39-71);
// transfers are the trilinear interpolation of the Q2 node lattice (DMDA Q1
// interpolation, src/ptatin_driver_linear_ts.c:996-1006) and its transpose;
// smoothing is the ChebychevRN recurrence (src/ksp_chebyrn.c:278-447) with the
// A11 diagonal as Jacobi preconditioner; the coarsest level is solved with a dense
// FP64 inverse.  All level work of one V-cycle is captured once into a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "ptq2.h"
#include "ptq2_ctx.h"
#include "ptq2_kernels.h"
#include "ptq2_solver.h"

using ptq2::Box;

namespace {

constexpr int TPB = 256;
// largest FGMRES restart length (Krylov pointer tables and reduction buffers are sized
// for it)
constexpr int MAXRESTART = 256;

inline unsigned nblk(int64_t n, int64_t cap = 16384) {
  int64_t b = (n + TPB - 1) / TPB;
  return (unsigned)std::max<int64_t>(1, std::min(b, cap));
}

// ------------------------------------------------------------ vector kernels
template <typename T>
__global__ void k_fill(int64_t n, T *x, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] = v;
}

// Input and output of the captured V-cycle: the graph's first kernel reads its input vector and
// scale, its last kernel its output vector, through this record, so one graph serves any
// caller vectors without copies (the record is set by k_set_io before each launch)
struct VcIO {
  const double *in;
  double *out;
  double sin, sout;
};
__global__ void k_set_io(VcIO *io, const double *in, double *out, double sin, double sout) {
  if (threadIdx.x == 0) {
    io->in = in;
    io->out = out;
    io->sin = sin;
    io->sout = sout;
  }
}
template <typename T>
__global__ void k_vin(const VcIO *__restrict__ io, int64_t n, T *__restrict__ b) {
  const double *in = io->in;
  const double s = io->sin;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = static_cast<T>(s * in[i]);
}
template <typename T>
__global__ void k_vout(const VcIO *__restrict__ io, int64_t n, const T *__restrict__ x) {
  double *out = io->out;
  const double s = io->sout;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = s * static_cast<double>(x[i]);
}


// p1 = scale * Dinv * r0 + x, with r0 = b (zero guess) or b - y
template <typename T>
__global__ void k_cheb_first(int64_t n, const T *__restrict__ b, const T *__restrict__ y, const T *__restrict__ dinv,
                             const T *__restrict__ x, T *__restrict__ p, T scale) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    T r = y ? b[i] - y[i] : b[i];
    T xi = x ? x[i] : T(0);
    p[i] = scale * dinv[i] * r + xi;
  }
}

// The Chebyshev updates fused with the face pass of the A11 apply before them: y holds the
// apply's owner-written values and face the upper-face slots of every brick (FACEVALS per
// brick, the layout k_gather_faces reads); a dof on a brick's lower face adds the slots of the
// lower neighbours in k_gather_faces' order, so the result is bitwise that of the unfused
// apply + update.  Dirichlet dofs take y = x (the apply input; MatMult semantics).  Then
//   out = cr * dinv * (b - y) + ck * pkm1 + cx * x       (pkm1 may be null: first step).
// The strided read-modify-write of the x faces that dominates the separate pass goes away.
// face_only: the apply's store already updated the Dirichlet dofs and every dof that no face
// slot adds to (store_tile_own<true>); only the lower-face dofs are left.
// four consecutive values through 16-byte accesses (the caller checks the alignment)
template <typename T>
__device__ __forceinline__ void ld4(const T *__restrict__ p, T v[4]) {
  if constexpr (sizeof(T) == 4) {
    const float4 a = *reinterpret_cast<const float4 *>(p);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  } else {
    const double2 a = reinterpret_cast<const double2 *>(p)[0], c = reinterpret_cast<const double2 *>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = c.x; v[3] = c.y;
  }
}
template <typename T>
__device__ __forceinline__ void st4(T *__restrict__ p, const T v[4]) {
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
  } else {
    reinterpret_cast<double2 *>(p)[0] = make_double2(v[0], v[1]);
    reinterpret_cast<double2 *>(p)[1] = make_double2(v[2], v[3]);
  }
}

template <typename T>
__global__ void k_cheb_faces(Box bx, int64_t n, const T *__restrict__ b, const T *__restrict__ y,
                             const T *__restrict__ dinv, const T *__restrict__ x, const T *__restrict__ pkm1,
                             T *__restrict__ out, const T *__restrict__ face, const uint8_t *__restrict__ mask, T cr,
                             T ck, T cx, int face_only) {
  const ptq2::ChebOut<T> co{b, dinv, x, pkm1, out, mask, cr, ck, cx};
  constexpr int TN = ptq2::TN, F2 = TN * TN, UP = 2 * ptq2::BE, NB = 2 * ptq2::BE;
  const unsigned NX = (unsigned)bx.NX, NXY = (unsigned)bx.NX * (unsigned)bx.NY;
  const int64_t sx = ptq2::FACEVALS, sy = (int64_t)bx.nbx * ptq2::FACEVALS,
                sz = (int64_t)bx.nbx * bx.nby * ptq2::FACEVALS;
  // the face slots added to the owner value v of dof i (a dof on a brick's lower face)
  auto faced = [&](int64_t i, T v) -> T {
    const unsigned nd = (unsigned)(i / 3), c = (unsigned)(i - 3 * (int64_t)nd);
    const unsigned Z = nd / NXY, rem = nd - Z * NXY, Yn = rem / NX, X = rem - Yn * NX;
    const int bxi = min((int)(X / NB), bx.nbx - 1), byi = min((int)(Yn / NB), bx.nby - 1),
              bzi = min((int)(Z / NB), bx.nbz - 1);
    const int lx = (int)X - NB * bxi, ly = (int)Yn - NB * byi, lz = (int)Z - NB * bzi;
    const bool nx = lx == 0 && bxi > 0, ny = ly == 0 && byi > 0, nz = lz == 0 && bzi > 0;
    if (!(nx || ny || nz)) return v;
    const T *fb = face + ((int64_t)bxi + (int64_t)bx.nbx * (byi + (int64_t)bx.nby * bzi)) * ptq2::FACEVALS;
    T acc = v;
    if (nz && ny && nx) acc += fb[-sx - sy - sz + 3 * (UP * TN + UP) + c];
    if (nz && ny) acc += fb[-sy - sz + 3 * (F2 + UP * TN + lx) + c];
    if (nz && nx) acc += fb[-sx - sz + 3 * (UP * TN + ly) + c];
    if (nz) acc += fb[-sz + 3 * (2 * F2 + ly * TN + lx) + c];
    if (ny && nx) acc += fb[-sx - sy + 3 * (lz * TN + UP) + c];
    if (ny) acc += fb[-sy + 3 * (F2 + lz * TN + lx) + c];
    if (nx) acc += fb[-sx + 3 * (lz * TN + ly) + c];
    return acc;
  };
  const bool al = (((uintptr_t)b | (uintptr_t)y | (uintptr_t)dinv | (uintptr_t)x | (uintptr_t)out |
                    (uintptr_t)(pkm1 ? pkm1 : b)) & 15) == 0 && ((uintptr_t)(mask ? mask : (const uint8_t *)b) & 3) == 0;
  int64_t i0 = 0;
  // FP32 levels of 2M dofs and more (64^3 elements and up): every dof is updated, four
  // consecutive dofs per step through 16-byte accesses.  Smaller levels keep one dof per thread
  // (with a quarter of the threads their latency shows), and the FP64 update gains nothing.
  if (sizeof(T) == 4 && !face_only && al && n >= (int64_t(1) << 21)) {
    const int64_t nq = n / 4;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += (int64_t)gridDim.x * blockDim.x) {
      const int64_t i = 4 * q;
      T yv[4], xv[4], bv[4], dv[4], pv[4] = {T(0), T(0), T(0), T(0)}, ov[4];
      ld4(y + i, yv);
      ld4(x + i, xv);
      ld4(b + i, bv);
      ld4(dinv + i, dv);
      if (pkm1) ld4(pkm1 + i, pv);
      const uint32_t mk = mask ? *reinterpret_cast<const uint32_t *>(mask + i) : 0u;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        T v;
        if ((mk >> (8 * t)) & 0xffu)
          v = xv[t];
        else
          v = face ? faced(i + t, yv[t]) : yv[t];
        ov[t] = ptq2::cheb_value(co, dv[t], bv[t], v, xv[t], pv[t], pkm1 != nullptr);
      }
      st4(out + i, ov);
    }
    i0 = 4 * nq;
  }
  for (int64_t i = i0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    T yv;
    if (mask && mask[i]) {
      if (face_only) continue;
      yv = x[i];
    } else {
      yv = y[i];
      if (face) {
        const unsigned nd = (unsigned)(i / 3), c = (unsigned)(i - 3 * (int64_t)nd);
        const unsigned Z = nd / NXY, rem = nd - Z * NXY, Yn = rem / NX, X = rem - Yn * NX;
        const int bxi = min((int)(X / NB), bx.nbx - 1), byi = min((int)(Yn / NB), bx.nby - 1),
                  bzi = min((int)(Z / NB), bx.nbz - 1);
        const int lx = (int)X - NB * bxi, ly = (int)Yn - NB * byi, lz = (int)Z - NB * bzi;
        const bool nx = lx == 0 && bxi > 0, ny = ly == 0 && byi > 0, nz = lz == 0 && bzi > 0;
        if (face_only && !(nx || ny || nz)) continue;
        if (nx || ny || nz) {
          const T *fb = face + ((int64_t)bxi + (int64_t)bx.nbx * (byi + (int64_t)bx.nby * bzi)) * ptq2::FACEVALS;
          T acc = yv;
          if (nz && ny && nx) acc += fb[-sx - sy - sz + 3 * (UP * TN + UP) + c];
          if (nz && ny) acc += fb[-sy - sz + 3 * (F2 + UP * TN + lx) + c];
          if (nz && nx) acc += fb[-sx - sz + 3 * (UP * TN + ly) + c];
          if (nz) acc += fb[-sz + 3 * (2 * F2 + ly * TN + lx) + c];
          if (ny && nx) acc += fb[-sx - sy + 3 * (lz * TN + UP) + c];
          if (ny) acc += fb[-sy + 3 * (F2 + lz * TN + lx) + c];
          if (nx) acc += fb[-sx + 3 * (lz * TN + ly) + c];
          yv = acc;
        }
      }
    }
    out[i] = ptq2::cheb_update(co, i, yv);
  }
}

// restriction R = P^T of the residual b - y onto the coarse node lattice, zero on
// coarse Dirichlet dofs.  One thread per coarse dof gathers its 3x3x3 fine stencil.
template <typename T>
__global__ void k_restrict(int NXc, int NYc, int NZc, int NXf, int NYf, int NZf, const T *__restrict__ b,
                           const T *__restrict__ y, const uint8_t *__restrict__ cmask, T *__restrict__ bc) {
  int64_t nc = 3 * (int64_t)NXc * NYc * NZc;
  for (int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; d < nc; d += (int64_t)gridDim.x * blockDim.x) {
    int comp = (int)(d % 3);
    int64_t node = d / 3;
    int I = (int)(node % NXc), J = (int)((node / NXc) % NYc), K = (int)(node / ((int64_t)NXc * NYc));
    T acc = T(0);
    if (!(cmask && cmask[d])) {
      for (int dk = -1; dk <= 1; dk++) {
        int k = 2 * K + dk;
        if (k < 0 || k >= NZf) continue;
        T wk = dk == 0 ? T(1) : T(0.5);
        for (int dj = -1; dj <= 1; dj++) {
          int j = 2 * J + dj;
          if (j < 0 || j >= NYf) continue;
          T wj = dj == 0 ? T(1) : T(0.5);
          for (int di = -1; di <= 1; di++) {
            int i = 2 * I + di;
            if (i < 0 || i >= NXf) continue;
            T wi = di == 0 ? T(1) : T(0.5);
            int64_t f = 3 * ((int64_t)i + (int64_t)NXf * ((int64_t)j + (int64_t)NYf * k)) + comp;
            T r = y ? b[f] - y[f] : b[f];
            acc += wk * wj * wi * r;
          }
        }
      }
    }
    bc[d] = acc;
  }
}

// x_f += P x_c (trilinear), not on fine Dirichlet dofs
template <typename T>
__global__ void k_prolong_add(int NXc, int NYc, int NXf, int NYf, int NZf, const T *__restrict__ xc,
                              const uint8_t *__restrict__ fmask, T *__restrict__ xf) {
  int64_t nf = 3 * (int64_t)NXf * NYf * NZf;
  for (int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; d < nf; d += (int64_t)gridDim.x * blockDim.x) {
    if (fmask && fmask[d]) continue;
    int comp = (int)(d % 3);
    int64_t node = d / 3;
    int i = (int)(node % NXf), j = (int)((node / NXf) % NYf), k = (int)(node / ((int64_t)NXf * NYf));
    int i0 = i >> 1, j0 = j >> 1, k0 = k >> 1;
    int i1 = (i & 1) ? i0 + 1 : i0, j1 = (j & 1) ? j0 + 1 : j0, k1 = (k & 1) ? k0 + 1 : k0;
    T wi = (i & 1) ? T(0.5) : T(1), wj = (j & 1) ? T(0.5) : T(1), wk = (k & 1) ? T(0.5) : T(1);
    auto at = [&](int a, int bb, int c) { return xc[3 * ((int64_t)a + (int64_t)NXc * ((int64_t)bb + (int64_t)NYc * c)) + comp]; };
    T v = at(i0, j0, k0);
    if (i & 1) v += at(i1, j0, k0);
    T s = v * wi;
    if (j & 1) {
      T v2 = at(i0, j1, k0);
      if (i & 1) v2 += at(i1, j1, k0);
      s = (s + v2 * wi) * wj;
    }
    if (k & 1) {
      T v3 = at(i0, j0, k1);
      if (i & 1) v3 += at(i1, j0, k1);
      T s3 = v3 * wi;
      if (j & 1) {
        T v4 = at(i0, j1, k1);
        if (i & 1) v4 += at(i1, j1, k1);
        s3 = (s3 + v4 * wi) * wj;
      }
      s = (s + s3) * wk;
    }
    xf[d] += s;
  }
}

// dense coarse solve x = Ainv b (FP64 inverse, one wave per row)
template <typename T>
__global__ void k_dense_gemv(int n, int ld, const double *__restrict__ Ainv, const T *__restrict__ b,
                             T *__restrict__ x) {
  int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  int lane = threadIdx.x & 63;
  if (row >= n) return;
  const double *a = Ainv + (int64_t)row * ld;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s += a[j] * (double)b[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) x[row] = static_cast<T>(s);
}

// coarse quadrature-point viscosity: the mean (arithmetic, geometric or harmonic) or the
// maximum of the fine quadrature values of the children that each coarse quadrature point touches (1, 2,
// 4 or 8 children).
__global__ void k_eta_coarsen(Box f, int mxc, int myc, int mzc, const double *__restrict__ eta_bm_f,
                              int mean, double *__restrict__ eta_c) {
  int64_t nq = 27 * (int64_t)mxc * myc * mzc;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nq; t += (int64_t)gridDim.x * blockDim.x) {
    int q = (int)(t % 27);
    int64_t e = t / 27;
    int cx = (int)(e % mxc), cy = (int)((e / mxc) % myc), cz = (int)(e / ((int64_t)mxc * myc));
    int qa = q % 3, qb = (q / 3) % 3, qc = q / 9;
    double s = 0.0;
    int cnt = 0;
    for (int c = 0; c < 2; c++) {
      if ((qc == 0 && c == 1) || (qc == 2 && c == 0)) continue;
      for (int bb = 0; bb < 2; bb++) {
        if ((qb == 0 && bb == 1) || (qb == 2 && bb == 0)) continue;
        for (int a = 0; a < 2; a++) {
          if ((qa == 0 && a == 1) || (qa == 2 && a == 0)) continue;
          int ex = 2 * cx + a, ey = 2 * cy + bb, ez = 2 * cz + c;
          int64_t lin = (ex >> 2) + (int64_t)f.nbx * ((ey >> 2) + (int64_t)f.nby * (ez >> 2));
          int lane = ptq2::elem_lane(ex & 3, ey & 3, ez & 3);
          const double *src = eta_bm_f + lin * 27 * ptq2::BEL + lane;
          for (int qq = 0; qq < 27; qq++) {
            const double v = src[qq * ptq2::BEL];
            if (mean == PTQ2_ETA_MEAN_MAX)
              s = fmax(s, v);
            else
              s += mean == PTQ2_ETA_MEAN_GEOMETRIC ? log(v) : (mean == PTQ2_ETA_MEAN_HARMONIC ? 1.0 / v : v);
          }
          cnt += 27;
        }
      }
    }
    const double a = mean == PTQ2_ETA_MEAN_MAX ? s : s / cnt;
    eta_c[t] = mean == PTQ2_ETA_MEAN_GEOMETRIC ? exp(a) : (mean == PTQ2_ETA_MEAN_HARMONIC ? 1.0 / a : a);
  }
}

// natural-order quadrature viscosity (e*27+q) from the brick-major copy
__global__ void k_eta_from_bm(Box f, const double *__restrict__ eta_bm, double *__restrict__ eta) {
  const int64_t nq = 27 * (int64_t)f.mx * f.my * f.mz;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nq; t += (int64_t)gridDim.x * blockDim.x) {
    const int q = (int)(t % 27);
    const int64_t e = t / 27;
    const int ex = (int)(e % f.mx), ey = (int)((e / f.mx) % f.my), ez = (int)(e / ((int64_t)f.mx * f.my));
    const int64_t lin = (ex >> 2) + (int64_t)f.nbx * ((ey >> 2) + (int64_t)f.nby * (ez >> 2));
    eta[t] = eta_bm[lin * 27 * ptq2::BEL + q * ptq2::BEL + ptq2::elem_lane(ex & 3, ey & 3, ez & 3)];
  }
}

// ---- the reference's coefficient hierarchy (SwarmUpdateGaussPropertiesLocalL2Projection_Q1_-
// MPntPStokes_Hierarchy, src/material_point_utils.c:1325-1524, projection type 0), on the host at
// set-up: a nodal field on the Q2 node lattice from the fine quadrature values, restricted lattice
// by lattice, interpolated to each coarse level's quadrature points.

// Q1 weight of corner c = cx + 2cy + 4cz at Gauss point p = i + 3j + 9k (element_utils_q1.c:32-47)
double q1_gauss_weight(int p, int c) {
  static const double g[3] = {-0.7745966692414834, 0.0, 0.7745966692414834};
  const int ix[3] = {p % 3, (p / 3) % 3, p / 9};
  double w = 0.125;
  for (int a = 0; a < 3; a++) w *= 1.0 + (((c >> a) & 1) ? g[ix[a]] : -g[ix[a]]);
  return w;
}

// _BuildQ1CoefficientProjection_QuadraturePoints_MPntPStokes_FineGrid (:1196-1300): per element and
// point, weight N_c(xi_p) on the 8 corner nodes and 1 on the others; node = sum(w eta) / sum(w)
void q1_nodal_projection(int mx, int my, int mz, const double *eta, std::vector<double> &nodal) {
  const int NX = 2 * mx + 1, NY = 2 * my + 1, NZ = 2 * mz + 1;
  std::vector<double> A((size_t)NX * NY * NZ, 0.0), B(A.size(), 0.0);
  double W[27][8], wsum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int p = 0; p < 27; p++)
    for (int c = 0; c < 8; c++) {
      W[p][c] = q1_gauss_weight(p, c);
      wsum[c] += W[p][c];
    }
  for (int ez = 0; ez < mz; ez++)
    for (int ey = 0; ey < my; ey++)
      for (int ex = 0; ex < mx; ex++) {
        const double *e = eta + 27 * ((int64_t)ex + (int64_t)mx * (ey + (int64_t)my * ez));
        double sum = 0.0, cw[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int p = 0; p < 27; p++) {
          sum += e[p];
          for (int c = 0; c < 8; c++) cw[c] += W[p][c] * e[p];
        }
        for (int k = 0; k < 3; k++)
          for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++) {
              const int64_t n = (int64_t)(2 * ex + i) + (int64_t)NX * ((2 * ey + j) + (int64_t)NY * (2 * ez + k));
              if ((i & 1) || (j & 1) || (k & 1)) {
                A[n] += sum;
                B[n] += 27.0;
              } else {
                const int c = (i >> 1) + 2 * (j >> 1) + 4 * (k >> 1);
                A[n] += cw[c];
                B[n] += wsum[c];
              }
            }
      }
  nodal.resize(A.size());
  for (size_t n = 0; n < A.size(); n++) nodal[n] = A[n] / B[n];
}

// MatRestrict with the trilinear lattice interpolation, scaled by 1 / (R 1): along each axis a
// fine node 2I weighs 1 and its odd neighbours 1/2, normalised by the weights present
void q1_restrict_nodal(int mxc, int myc, int mzc, const std::vector<double> &f, std::vector<double> &c) {
  const int NXc = 2 * mxc + 1, NYc = 2 * myc + 1, NZc = 2 * mzc + 1;
  const int NXf = 2 * NXc - 1, NYf = 2 * NYc - 1;
  c.assign((size_t)NXc * NYc * NZc, 0.0);
  for (int K = 0; K < NZc; K++)
    for (int J = 0; J < NYc; J++)
      for (int I = 0; I < NXc; I++) {
        double num = 0.0, den = 0.0;
        for (int dk = -1; dk <= 1; dk++) {
          const int fk = 2 * K + dk;
          if (fk < 0 || fk > 2 * (NZc - 1)) continue;
          const double wk = dk ? 0.5 : 1.0;
          for (int dj = -1; dj <= 1; dj++) {
            const int fj = 2 * J + dj;
            if (fj < 0 || fj > 2 * (NYc - 1)) continue;
            const double wj = dj ? 0.5 : 1.0;
            for (int di = -1; di <= 1; di++) {
              const int fi = 2 * I + di;
              if (fi < 0 || fi > 2 * (NXc - 1)) continue;
              const double w = wk * wj * (di ? 0.5 : 1.0);
              num += w * f[(size_t)fi + (size_t)NXf * (fj + (size_t)NYf * fk)];
              den += w;
            }
          }
        }
        c[(size_t)I + (size_t)NXc * (J + (size_t)NYc * K)] = num / den;
      }
}

// _SwarmUpdateGaussPropertiesLocalL2ProjectionQ1_MPntPStokes_InterpolateToQuadratePoints
// (:957-1070): Q1 interpolation of the corner nodes, clipped to the element's nodal range
void q1_interpolate_qp(int mx, int my, int mz, const std::vector<double> &nodal, std::vector<double> &eta) {
  const int NX = 2 * mx + 1, NY = 2 * my + 1;
  eta.resize((size_t)27 * mx * my * mz);
  for (int ez = 0; ez < mz; ez++)
    for (int ey = 0; ey < my; ey++)
      for (int ex = 0; ex < mx; ex++) {
        double cv[8], lo = 1e300, hi = -1e300;
        for (int k = 0; k < 3; k++)
          for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++) {
              const double v = nodal[(size_t)(2 * ex + i) + (size_t)NX * ((2 * ey + j) + (size_t)NY * (2 * ez + k))];
              lo = std::min(lo, v);
              hi = std::max(hi, v);
              if (!((i & 1) || (j & 1) || (k & 1))) cv[(i >> 1) + 2 * (j >> 1) + 4 * (k >> 1)] = v;
            }
        double *out = &eta[27 * ((size_t)ex + (size_t)mx * (ey + (size_t)my * ez))];
        for (int p = 0; p < 27; p++) {
          double s = 0.0;
          for (int c = 0; c < 8; c++) s += q1_gauss_weight(p, c) * cv[c];
          out[p] = std::min(std::max(s, lo), hi);
        }
      }
}

// Dot products and updates of the Gram-Schmidt passes.  The reductions run on a fixed
// grid of MDB blocks: every block writes one partial per product and k_sum_parts adds a
// product's partials in block order, so every sum is reproducible run to run.
constexpr int MD = 8;      // vectors per pass over w
constexpr int MDB = 2048;  // blocks of the reduction kernels (partials per product)
constexpr int MDU = 4;     // grid strides per loop iteration (independent loads in flight)

template <int NR>
__device__ __forceinline__ void block_partials(double (&s)[NR], double (*red)[TPB / 64]) {
#pragma unroll
  for (int m = 0; m < NR; m++) {
    for (int off = 32; off > 0; off >>= 1) s[m] += __shfl_down(s[m], off, 64);
    if ((threadIdx.x & 63) == 0) red[m][threadIdx.x / 64] = s[m];
  }
}

// part[v * MDB + block]: partial of V[v].w for v < nv; with self, part[nv * MDB + block]:
// partial of w.w (taken in the first pass over w)
template <typename VT>
__global__ __launch_bounds__(TPB) void k_mdot(int64_t n, const VT *const *__restrict__ V, int nv,
                                              const double *__restrict__ w, int self, double *__restrict__ part) {
  __shared__ double red[MD + 1][TPB / 64];
  const int64_t stride = (int64_t)MDB * TPB;
  const int64_t i0 = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const int ngroups = nv > 0 ? (nv + MD - 1) / MD : 1;
  for (int gi = 0; gi < ngroups; gi++) {
    const int v0 = gi * MD;
    const int nm = max(0, min(MD, nv - v0));
    const bool ws = self && gi == 0;
    double s[MD + 1];
    const VT *vp[MD];
#pragma unroll
    for (int m = 0; m <= MD; m++) s[m] = 0.0;
#pragma unroll
    for (int m = 0; m < MD; m++) vp[m] = m < nm ? V[v0 + m] : nullptr;  // read only for m < nm
    int64_t i = i0;
    for (; i + (MDU - 1) * stride < n; i += MDU * stride) {
      double wv[MDU];
#pragma unroll
      for (int u = 0; u < MDU; u++) wv[u] = w[i + u * stride];
#pragma unroll
      for (int m = 0; m < MD; m++)
        if (m < nm) {
          double x[MDU];
#pragma unroll
          for (int u = 0; u < MDU; u++) x[u] = (double)vp[m][i + u * stride];
#pragma unroll
          for (int u = 0; u < MDU; u++) s[m] = fma(x[u], wv[u], s[m]);
        }
      if (ws)
#pragma unroll
        for (int u = 0; u < MDU; u++) s[MD] = fma(wv[u], wv[u], s[MD]);
    }
    for (; i < n; i += stride) {
      const double wi = w[i];
#pragma unroll
      for (int m = 0; m < MD; m++)
        if (m < nm) s[m] = fma((double)vp[m][i], wi, s[m]);
      if (ws) s[MD] = fma(wi, wi, s[MD]);
    }
    block_partials<MD + 1>(s, red);
    __syncthreads();
    const int t = threadIdx.x;
    if (t < nm || (ws && t == MD)) {
      const double r = red[t][0] + red[t][1] + red[t][2] + red[t][3];
      part[(int64_t)(t < nm ? v0 + t : nv) * MDB + blockIdx.x] = r;
    }
    __syncthreads();
  }
}

// out[p] = sum of the MDB partials of product p, in block order (one block per product)
__global__ __launch_bounds__(TPB) void k_sum_parts(const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double red[1][TPB / 64];
  const double *p = part + (int64_t)blockIdx.x * MDB;
  double s[1] = {0.0};
  for (int k = threadIdx.x; k < MDB; k += TPB) s[0] += p[k];
  block_partials<1>(s, red);
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
}

// w -= sum_v h[v] V[v], and part[block]: partial of the updated w.w
template <typename VT>
__global__ __launch_bounds__(TPB) void k_maxpy_norm(int64_t n, const VT *const *__restrict__ V, int nv,
                                                    const double *__restrict__ h, double *__restrict__ w,
                                                    double *__restrict__ part) {
  __shared__ double red[1][TPB / 64];
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)MDB * TPB) {
    double s = w[i];
    for (int v = 0; v < nv; v++) s -= h[v] * (double)V[v][i];
    w[i] = s;
    acc[0] = fma(s, s, acc[0]);
  }
  block_partials<1>(acc, red);
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
}

// Gram-Schmidt with delayed reorthogonalisation (CGS2 in two passes per Arnoldi step): the
// pending candidate u (stored unnormalised, u = su * slot) is reorthogonalised against the
// finished basis V in the same pass that projects the new vector w.
// Pass 1, part rows: 2v, 2v+1 -> V[v].u', V[v].w (v < nv); 2nv, 2nv+1, 2nv+2 -> u'.w, u'.u',
// w.w, with u' the stored (unscaled) candidate.  MD2 basis vectors per pass over u' and w.
constexpr int MD2 = 16;
template <typename VT>
__global__ __launch_bounds__(TPB) void k_mdot2(int64_t n, const VT *const *__restrict__ V, int nv,
                                               const double *__restrict__ u, const double *__restrict__ w,
                                               double *__restrict__ part) {
  constexpr int NR = 2 * MD2 + 3;
  __shared__ double red[NR][TPB / 64];
  const int64_t stride = (int64_t)MDB * TPB;
  const int64_t i0 = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const int ngroups = nv > 0 ? (nv + MD2 - 1) / MD2 : 1;
  for (int gi = 0; gi < ngroups; gi++) {
    const int v0 = gi * MD2;
    const int nm = max(0, min(MD2, nv - v0));
    const bool first = gi == 0;
    double s[NR];
    const VT *vp[MD2];
#pragma unroll
    for (int m = 0; m < NR; m++) s[m] = 0.0;
#pragma unroll
    for (int m = 0; m < MD2; m++) vp[m] = m < nm ? V[v0 + m] : nullptr;  // read only for m < nm
    int64_t i = i0;
    for (; i + (MDU - 1) * stride < n; i += MDU * stride) {
      double uv[MDU], wv[MDU];
#pragma unroll
      for (int q = 0; q < MDU; q++) {
        uv[q] = u[i + q * stride];
        wv[q] = w[i + q * stride];
      }
#pragma unroll
      for (int m = 0; m < MD2; m++)
        if (m < nm) {
          double x[MDU];
#pragma unroll
          for (int q = 0; q < MDU; q++) x[q] = (double)vp[m][i + q * stride];
#pragma unroll
          for (int q = 0; q < MDU; q++) {
            s[2 * m] = fma(x[q], uv[q], s[2 * m]);
            s[2 * m + 1] = fma(x[q], wv[q], s[2 * m + 1]);
          }
        }
      if (first)
#pragma unroll
        for (int q = 0; q < MDU; q++) {
          s[2 * MD2] = fma(uv[q], wv[q], s[2 * MD2]);
          s[2 * MD2 + 1] = fma(uv[q], uv[q], s[2 * MD2 + 1]);
          s[2 * MD2 + 2] = fma(wv[q], wv[q], s[2 * MD2 + 2]);
        }
    }
    for (; i < n; i += stride) {
      const double ui = u[i], wi = w[i];
#pragma unroll
      for (int m = 0; m < MD2; m++)
        if (m < nm) {
          const double x = (double)vp[m][i];
          s[2 * m] = fma(x, ui, s[2 * m]);
          s[2 * m + 1] = fma(x, wi, s[2 * m + 1]);
        }
      if (first) {
        s[2 * MD2] = fma(ui, wi, s[2 * MD2]);
        s[2 * MD2 + 1] = fma(ui, ui, s[2 * MD2 + 1]);
        s[2 * MD2 + 2] = fma(wi, wi, s[2 * MD2 + 2]);
      }
    }
    block_partials<NR>(s, red);
    __syncthreads();
    const int t = threadIdx.x;
    int row = -1;
    if (t < 2 * nm) row = 2 * v0 + t;
    else if (first && t >= 2 * MD2 && t < NR) row = 2 * nv + (t - 2 * MD2);
    if (row >= 0) part[(int64_t)row * MDB + blockIdx.x] = red[t][0] + red[t][1] + red[t][2] + red[t][3];
    __syncthreads();
  }
}

// Pass 2: with uu = su * u' and the coefficient table ab = [a (nv), b (nv)]:
//   vo <- (uu - sum_v a[v] V[v]) * ia      (the finished basis vector; vo may be u' itself)
//   wo <- w - sum_v b[v] V[v] - gam * uu   (the next candidate, unnormalised)
// and part[block]: partial of wo.wo.
template <typename VT>
__global__ __launch_bounds__(TPB) void k_dcgs_update(int64_t n, const VT *const *__restrict__ V, int nv,
                                                     const double *__restrict__ ab, double su, double ia, double gam,
                                                     const double *u, VT *vo, const double *__restrict__ w,
                                                     double *__restrict__ wo, double *__restrict__ part) {
  __shared__ double red[1][TPB / 64];
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)MDB * TPB) {
    double sa = 0.0, sb = 0.0;
    for (int v = 0; v < nv; v++) {
      const double x = (double)V[v][i];
      sa = fma(ab[v], x, sa);
      sb = fma(ab[nv + v], x, sb);
    }
    const double uu = su * u[i];
    vo[i] = (VT)((uu - sa) * ia);
    const double r = w[i] - sb - gam * uu;
    wo[i] = r;
    acc[0] = fma(r, r, acc[0]);
  }
  block_partials<1>(acc, red);
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
}

// FP32-basis forms of the two passes: each thread takes element pairs (float2 of the basis,
// double2 of u' and w), so a wave moves as many bytes per load as the FP64 forms; an odd last
// element is added by thread 0 of block 0.  Same partial layout as k_mdot2 / k_dcgs_update.
__global__ __launch_bounds__(TPB) void k_mdot2_f32(int64_t n, const float *const *__restrict__ V, int nv,
                                                   const double *__restrict__ u, const double *__restrict__ w,
                                                   double *__restrict__ part) {
  constexpr int NR = 2 * MD2 + 3;
  __shared__ double red[NR][TPB / 64];
  const int64_t np = n >> 1;
  const int64_t stride = (int64_t)MDB * TPB;
  const int64_t i0 = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const bool tail = (n & 1) && i0 == 0;
  const double2 *u2 = (const double2 *)u, *w2 = (const double2 *)w;
  const int ngroups = nv > 0 ? (nv + MD2 - 1) / MD2 : 1;
  for (int gi = 0; gi < ngroups; gi++) {
    const int v0 = gi * MD2;
    const int nm = max(0, min(MD2, nv - v0));
    const bool first = gi == 0;
    double s[NR];
    const float2 *vp[MD2];
#pragma unroll
    for (int m = 0; m < NR; m++) s[m] = 0.0;
#pragma unroll
    for (int m = 0; m < MD2; m++) vp[m] = m < nm ? (const float2 *)V[v0 + m] : nullptr;  // read only for m < nm
    int64_t i = i0;
    for (; i + (MDU - 1) * stride < np; i += MDU * stride) {
      double2 uv[MDU], wv[MDU];
#pragma unroll
      for (int q = 0; q < MDU; q++) {
        uv[q] = u2[i + q * stride];
        wv[q] = w2[i + q * stride];
      }
#pragma unroll
      for (int m = 0; m < MD2; m++)
        if (m < nm) {
          float2 x[MDU];
#pragma unroll
          for (int q = 0; q < MDU; q++) x[q] = vp[m][i + q * stride];
#pragma unroll
          for (int q = 0; q < MDU; q++) {
            s[2 * m] = fma((double)x[q].x, uv[q].x, s[2 * m]);
            s[2 * m] = fma((double)x[q].y, uv[q].y, s[2 * m]);
            s[2 * m + 1] = fma((double)x[q].x, wv[q].x, s[2 * m + 1]);
            s[2 * m + 1] = fma((double)x[q].y, wv[q].y, s[2 * m + 1]);
          }
        }
      if (first)
#pragma unroll
        for (int q = 0; q < MDU; q++) {
          s[2 * MD2] = fma(uv[q].x, wv[q].x, s[2 * MD2]);
          s[2 * MD2] = fma(uv[q].y, wv[q].y, s[2 * MD2]);
          s[2 * MD2 + 1] = fma(uv[q].x, uv[q].x, s[2 * MD2 + 1]);
          s[2 * MD2 + 1] = fma(uv[q].y, uv[q].y, s[2 * MD2 + 1]);
          s[2 * MD2 + 2] = fma(wv[q].x, wv[q].x, s[2 * MD2 + 2]);
          s[2 * MD2 + 2] = fma(wv[q].y, wv[q].y, s[2 * MD2 + 2]);
        }
    }
    for (; i < np; i += stride) {
      const double2 ui = u2[i], wi = w2[i];
#pragma unroll
      for (int m = 0; m < MD2; m++)
        if (m < nm) {
          const float2 x = vp[m][i];
          s[2 * m] = fma((double)x.y, ui.y, fma((double)x.x, ui.x, s[2 * m]));
          s[2 * m + 1] = fma((double)x.y, wi.y, fma((double)x.x, wi.x, s[2 * m + 1]));
        }
      if (first) {
        s[2 * MD2] = fma(ui.y, wi.y, fma(ui.x, wi.x, s[2 * MD2]));
        s[2 * MD2 + 1] = fma(ui.y, ui.y, fma(ui.x, ui.x, s[2 * MD2 + 1]));
        s[2 * MD2 + 2] = fma(wi.y, wi.y, fma(wi.x, wi.x, s[2 * MD2 + 2]));
      }
    }
    if (tail) {
      const int64_t e = n - 1;
#pragma unroll
      for (int m = 0; m < MD2; m++)
        if (m < nm) {
          const double x = (double)V[v0 + m][e];
          s[2 * m] = fma(x, u[e], s[2 * m]);
          s[2 * m + 1] = fma(x, w[e], s[2 * m + 1]);
        }
      if (first) {
        s[2 * MD2] = fma(u[e], w[e], s[2 * MD2]);
        s[2 * MD2 + 1] = fma(u[e], u[e], s[2 * MD2 + 1]);
        s[2 * MD2 + 2] = fma(w[e], w[e], s[2 * MD2 + 2]);
      }
    }
    block_partials<NR>(s, red);
    __syncthreads();
    const int t = threadIdx.x;
    int row = -1;
    if (t < 2 * nm) row = 2 * v0 + t;
    else if (first && t >= 2 * MD2 && t < NR) row = 2 * nv + (t - 2 * MD2);
    if (row >= 0) part[(int64_t)row * MDB + blockIdx.x] = red[t][0] + red[t][1] + red[t][2] + red[t][3];
    __syncthreads();
  }
}

__global__ __launch_bounds__(TPB) void k_dcgs_update_f32(int64_t n, const float *const *__restrict__ V, int nv,
                                                         const double *__restrict__ ab, double su, double ia,
                                                         double gam, const double *__restrict__ u,
                                                         float *__restrict__ vo, const double *__restrict__ w,
                                                         double *__restrict__ wo, double *__restrict__ part) {
  __shared__ double red[1][TPB / 64];
  double acc[1] = {0.0};
  const int64_t np = n >> 1;
  const double2 *u2 = (const double2 *)u, *w2 = (const double2 *)w;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < np; i += (int64_t)MDB * TPB) {
    double sa0 = 0.0, sa1 = 0.0, sb0 = 0.0, sb1 = 0.0;
    for (int v = 0; v < nv; v++) {
      const float2 x = ((const float2 *)V[v])[i];
      sa0 = fma(ab[v], (double)x.x, sa0);
      sa1 = fma(ab[v], (double)x.y, sa1);
      sb0 = fma(ab[nv + v], (double)x.x, sb0);
      sb1 = fma(ab[nv + v], (double)x.y, sb1);
    }
    const double2 ui = u2[i], wi = w2[i];
    const double uu0 = su * ui.x, uu1 = su * ui.y;
    ((float2 *)vo)[i] = make_float2((float)((uu0 - sa0) * ia), (float)((uu1 - sa1) * ia));
    const double r0 = wi.x - sb0 - gam * uu0, r1 = wi.y - sb1 - gam * uu1;
    ((double2 *)wo)[i] = make_double2(r0, r1);
    acc[0] = fma(r1, r1, fma(r0, r0, acc[0]));
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const int64_t e = n - 1;
    double sa = 0.0, sb = 0.0;
    for (int v = 0; v < nv; v++) {
      const double x = (double)V[v][e];
      sa = fma(ab[v], x, sa);
      sb = fma(ab[nv + v], x, sb);
    }
    const double uu = su * u[e];
    vo[e] = (float)((uu - sa) * ia);
    const double r = w[e] - sb - gam * uu;
    wo[e] = r;
    acc[0] = fma(r, r, acc[0]);
  }
  block_partials<1>(acc, red);
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
}

// x += sum_v h[v] V[v]
__global__ void k_mcombine(int64_t n, const double *const *__restrict__ V, int nv, const double *__restrict__ h,
                           double *__restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double s = x[i];
    for (int v = 0; v < nv; v++) s += h[v] * V[v][i];
    x[i] = s;
  }
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

__global__ void k_scale_copy(int64_t n, const double *__restrict__ a, double s, double *__restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = s * a[i];
}

__global__ void k_sub(int64_t n, const double *__restrict__ b, const double *__restrict__ y, double *__restrict__ r) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    r[i] = b[i] - y[i];
}

// partial sums of v.w and v.(d v) per block: part[2*block], part[2*block+1]
__global__ void k_rq_parts(int64_t n, const double *__restrict__ v, const double *__restrict__ w,
                           const double *__restrict__ d, double *__restrict__ part) {
  __shared__ double r0[TPB / 64], r1[TPB / 64];
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    a += v[i] * w[i];
    b += v[i] * d[i] * v[i];
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    r0[threadIdx.x / 64] = a;
    r1[threadIdx.x / 64] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0.0, t1 = 0.0;
    for (int k = 0; k < TPB / 64; k++) {
      t0 += r0[k];
      t1 += r1[k];
    }
    part[2 * blockIdx.x] = t0;
    part[2 * blockIdx.x + 1] = t1;
  }
}

__global__ void k_div(int64_t n, const double *__restrict__ w, const double *__restrict__ d, double *__restrict__ v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = w[i] / d[i];
}

template <typename T>
__global__ void k_recip(int64_t n, const double *__restrict__ d, T *__restrict__ dinv) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dinv[i] = static_cast<T>(1.0 / d[i]);
}

// ------------------------------------------------------------------- levels

// ---------------------------------------------------------------- rigid-body deflation
// Opt-in two-level correction for stiff inclusions (ptq2_mg_opts.rigid_deflation; the
// reference's velocity solve is PCMG + FGMRES only and has no counterpart).  An inclusion is a
// connected set of elements (sharing a node) that hold a quadrature viscosity above
// sqrt(min * max) of the field; its 6 rigid motions -- translations e_a and rotations
// e_a x (x - c) / L on the nodes of its elements (c the mean of those nodes, L their RMS
// distance to it), Dirichlet dofs zeroed -- are columns of W.  Around the V-cycle M^-1 the
// preconditioner becomes (A-DEF2 of Tang, Nabben, Vuik and Erlangga, J. Sci. Comput. 39, 2009)
//   z = W y + M^-1 (r - A W y),   y = E^-1 W^T r,   E = W^T A W.
// A W is kept on the nodes it reaches (the inclusion's elements and their neighbours), so the
// correction needs no operator apply: W^T r as chunked moments over the inclusions' nodes,
// y by a dense product with E^-1, then two updates over those nodes.  Fixed summation order
// throughout.
constexpr int RG_CH = 1024;     // inclusion nodes per moment block
constexpr int RG_MAXINC = 128;  // inclusions (6 columns each)

struct RgChunk {
  int comp, begin, end, pad;
};

// part[6 q + k]: moment k of x over chunk q (k < 3: sum of x_a; 3..5: sum of d x x, d the
// scaled offset from the inclusion's centre), Dirichlet dofs excluded
template <typename T>
__global__ __launch_bounds__(TPB) void k_rg_moments(const T *__restrict__ x, const uint8_t *__restrict__ mask,
                                                   const double *__restrict__ coords, const int *__restrict__ nodes,
                                                   const RgChunk *__restrict__ chunk, const double4 *__restrict__ cent,
                                                   double *__restrict__ part) {
  __shared__ double red[6][TPB / 64];
  const RgChunk ch = chunk[blockIdx.x];
  const double4 g = cent[ch.comp];
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = ch.begin + (int)threadIdx.x; i < ch.end; i += TPB) {
    const int64_t o = 3 * (int64_t)nodes[i];
    double r[3];
#pragma unroll
    for (int a = 0; a < 3; a++) r[a] = (mask && mask[o + a]) ? 0.0 : (double)x[o + a];
    const double d0 = (coords[o] - g.x) * g.w, d1 = (coords[o + 1] - g.y) * g.w, d2 = (coords[o + 2] - g.z) * g.w;
    s[0] += r[0];
    s[1] += r[1];
    s[2] += r[2];
    s[3] += d1 * r[2] - d2 * r[1];
    s[4] += d2 * r[0] - d0 * r[2];
    s[5] += d0 * r[1] - d1 * r[0];
  }
  block_partials<6>(s, red);
  __syncthreads();
  if (threadIdx.x < 6) {
    const int m = threadIdx.x;
    part[6 * (int64_t)blockIdx.x + m] = red[m][0] + red[m][1] + red[m][2] + red[m][3];
  }
}

// m[6 c + k] = moment k of inclusion c: its chunk partials summed by one block per inclusion
// (fixed strides and tree, so reproducible); set-up: m is a column of E
__global__ __launch_bounds__(TPB) void k_rg_msum(const int *__restrict__ cchunk, const double *__restrict__ part,
                                                double *__restrict__ m) {
  __shared__ double red[6][TPB / 64];
  const int c = blockIdx.x;
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int q = cchunk[c] + (int)threadIdx.x; q < cchunk[c + 1]; q += TPB)
#pragma unroll
    for (int k = 0; k < 6; k++) s[k] += part[6 * (int64_t)q + k];
  block_partials<6>(s, red);
  __syncthreads();
  if (threadIdx.x < 6) m[6 * c + threadIdx.x] = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] +
                                                red[threadIdx.x][3];
}

// y = E^-1 m; 4 rows per block, one wave per row
__global__ __launch_bounds__(256) void k_rg_coarse(int n6, const double *__restrict__ m,
                                                  const double *__restrict__ einv, double *__restrict__ y) {
  const int row = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n6) return;  // whole waves leave together
  double s = 0.0;
  for (int j = lane; j < n6; j += 64) s += einv[(int64_t)row * n6 + j] * m[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[row] = s;
}

// b -= (A W) y on the nodes A W reaches
template <typename T>
__global__ void k_rg_subtract(int64_t nd, const int *__restrict__ dnode, const int *__restrict__ dptr,
                              const int *__restrict__ ecomp, const double *__restrict__ aw,
                              const double *__restrict__ y, T *__restrict__ b) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nd; t += (int64_t)gridDim.x * blockDim.x) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int e = dptr[t]; e < dptr[t + 1]; e++) {
      const double *yc = y + 6 * ecomp[e];
      const double *w = aw + 18 * (int64_t)e;
#pragma unroll
      for (int k = 0; k < 6; k++) {
        a0 += w[3 * k] * yc[k];
        a1 += w[3 * k + 1] * yc[k];
        a2 += w[3 * k + 2] * yc[k];
      }
    }
    const int64_t o = 3 * (int64_t)dnode[t];
    b[o] = (T)((double)b[o] - a0);
    b[o + 1] = (T)((double)b[o + 1] - a1);
    b[o + 2] = (T)((double)b[o + 2] - a2);
  }
}

// x += W y on the inclusions' nodes (Dirichlet dofs untouched)
template <typename T>
__global__ __launch_bounds__(TPB) void k_rg_expand(const uint8_t *__restrict__ mask, const double *__restrict__ coords,
                                                  const int *__restrict__ nodes, const RgChunk *__restrict__ chunk,
                                                  const double4 *__restrict__ cent, const double *__restrict__ y,
                                                  T *__restrict__ x) {
  const RgChunk ch = chunk[blockIdx.x];
  const double4 g = cent[ch.comp];
  const double *yc = y + 6 * ch.comp;
  const double t0 = yc[0], t1 = yc[1], t2 = yc[2], w0 = yc[3], w1 = yc[4], w2 = yc[5];
  for (int i = ch.begin + (int)threadIdx.x; i < ch.end; i += TPB) {
    const int64_t o = 3 * (int64_t)nodes[i];
    const double d0 = (coords[o] - g.x) * g.w, d1 = (coords[o + 1] - g.y) * g.w, d2 = (coords[o + 2] - g.z) * g.w;
    const double u[3] = {t0 + w1 * d2 - w2 * d1, t1 + w2 * d0 - w0 * d2, t2 + w0 * d1 - w1 * d0};
#pragma unroll
    for (int a = 0; a < 3; a++)
      if (!(mask && mask[o + a])) x[o + a] = (T)((double)x[o + a] + u[a]);
  }
}

// the range of the quadrature viscosity of every element (natural element order)
__global__ void k_eta_elem_range(Box f, const double *__restrict__ eta_bm, double *__restrict__ emax,
                                 double *__restrict__ emin) {
  const int64_t ne = (int64_t)f.mx * f.my * f.mz;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    const int ex = (int)(e % f.mx), ey = (int)((e / f.mx) % f.my), ez = (int)(e / ((int64_t)f.mx * f.my));
    const int64_t lin = (ex >> 2) + (int64_t)f.nbx * ((ey >> 2) + (int64_t)f.nby * (ez >> 2));
    const double *src = eta_bm + lin * 27 * ptq2::BEL + ptq2::elem_lane(ex & 3, ey & 3, ez & 3);
    double hi = src[0], lo = src[0];
    for (int q = 1; q < 27; q++) {
      hi = fmax(hi, src[q * ptq2::BEL]);
      lo = fmin(lo, src[q * ptq2::BEL]);
    }
    emax[e] = hi;
    emin[e] = lo;
  }
}

// aw[18 e + 3 k + a] = v[3 node + a] for the entries e of inclusion c (column k of its A W)
__global__ void k_rg_gather(int64_t nd, const int *__restrict__ dnode, const int *__restrict__ dptr,
                            const int *__restrict__ ecomp, int c, int k, const double *__restrict__ v,
                            double *__restrict__ aw) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nd; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = 3 * (int64_t)dnode[t];
    for (int e = dptr[t]; e < dptr[t + 1]; e++)
      if (ecomp[e] == c)
        for (int a = 0; a < 3; a++) aw[18 * (int64_t)e + 3 * k + a] = v[o + a];
  }
}

// column j of the operator assembled into row j of A (ld columns): e = unit vector j
__global__ void k_unit(int64_t n, int64_t j, double *__restrict__ e) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) e[i] = i == j ? 1.0 : 0.0;
}
__global__ void k_pad_identity(int nc, int np, double *__restrict__ A) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)np * np) return;
  int i = (int)(idx / np), j = (int)(idx % np);
  if (i >= nc || j >= nc) A[idx] = i == j ? 1.0 : 0.0;
}
// rows of coarse Dirichlet dofs: the unit row (the MatMult row y_D = x_D)
__global__ void k_dirichlet_unit_rows(int64_t n, const uint8_t *__restrict__ mask, double *__restrict__ A, int ld) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && mask && mask[i]) A[i * ld + i] = 1.0;
}
// the dof of the probed colour class (node indices = (cx,cy,cz) modulo pp on each axis,
// component comp) nearest to node (I,J,K): offsets -pr..pr per axis, -1 if outside the box
__device__ __forceinline__ int64_t probe_member(int NX, int NY, int NZ, int pp, int cx, int cy, int cz, int comp,
                                                int I, int J, int K) {
  const int pr = pp >> 1;
  int d = (cx - I % pp + pp) % pp, Ij = I + (d <= pr ? d : d - pp);
  d = (cy - J % pp + pp) % pp;
  int Jj = J + (d <= pr ? d : d - pp);
  d = (cz - K % pp + pp) % pp;
  int Kj = K + (d <= pr ? d : d - pp);
  if (Ij < 0 || Ij >= NX || Jj < 0 || Jj >= NY || Kj < 0 || Kj >= NZ) return -1;
  return 3 * ((int64_t)Ij + (int64_t)NX * ((int64_t)Jj + (int64_t)NY * Kj)) + comp;
}
// probe vector of a colour class: 1 on its non-Dirichlet dofs, 0 elsewhere
__global__ void k_probe_vec(int NX, int NY, int NZ, int pp, int cx, int cy, int cz, int comp,
                            const uint8_t *__restrict__ mask, double *__restrict__ e) {
  const int64_t n = 3 * (int64_t)NX * NY * NZ;
  for (int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; d < n; d += (int64_t)gridDim.x * blockDim.x) {
    const int64_t node = d / 3;
    const int I = (int)(node % NX), J = (int)((node / NX) % NY), K = (int)(node / ((int64_t)NX * NY));
    const bool on = (int)(d - 3 * node) == comp && I % pp == cx && J % pp == cy && K % pp == cz;
    e[d] = on && !(mask && mask[d]) ? 1.0 : 0.0;
  }
}
// entry r of the probe result belongs to the column of the class member next to dof r's node:
// row j of A holds column j of the operator (A[j][r] = (A e_j)[r])
__global__ void k_probe_scatter(int NX, int NY, int NZ, int pp, int cx, int cy, int cz, int comp,
                                const uint8_t *__restrict__ mask, const double *__restrict__ w, double *__restrict__ A,
                                int ld) {
  const int64_t n = 3 * (int64_t)NX * NY * NZ;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t node = r / 3;
    const int I = (int)(node % NX), J = (int)((node / NX) % NY), K = (int)(node / ((int64_t)NX * NY));
    const int64_t j = probe_member(NX, NY, NZ, pp, cx, cy, cz, comp, I, J, K);
    if (j >= 0 && !(mask && mask[j])) A[j * ld + r] = w[r];
  }
}
// Galerkin levels as assembled stencils.  The Galerkin product of the trilinear transfers
// around a Q2 operator couples a node to the nodes within two nodes per axis, so a level's
// operator is a 5x5x5 stencil of 3x3 blocks per node: A[((o * 3 + a) * 3 + b) * nn + p] couples
// component a of node p to component b of node p + offset(o), o = (dz+2) 25 + (dy+2) 5 + (dx+2)
// (node-fastest, so a thread per node reads it coalesced).  Entries towards nodes outside the
// box are zero.
constexpr int SN = 125, SC = 62;  // stencil offsets, the centre
// entry (p, a) of a probe result (colour class (cx,cy,cz) of period 5, component comp) is the
// coupling of p to the class member next to it
__global__ void k_stencil_scatter(int NX, int NY, int NZ, int cx, int cy, int cz, int comp,
                                  const uint8_t *__restrict__ mask, const double *__restrict__ w,
                                  double *__restrict__ A) {
  const int64_t nn = (int64_t)NX * NY * NZ;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nn; p += (int64_t)gridDim.x * blockDim.x) {
    const int I = (int)(p % NX), J = (int)((p / NX) % NY), K = (int)(p / ((int64_t)NX * NY));
    const int64_t q = probe_member(NX, NY, NZ, 5, cx, cy, cz, comp, I, J, K);
    if (q < 0 || (mask && mask[q])) continue;
    const int64_t qn = q / 3;
    const int Iq = (int)(qn % NX), Jq = (int)((qn / NX) % NY), Kq = (int)(qn / ((int64_t)NX * NY));
    const int o = (Kq - K + 2) * 25 + (Jq - J + 2) * 5 + (Iq - I + 2);
#pragma unroll
    for (int a = 0; a < 3; a++) A[((int64_t)(o * 3 + a) * 3 + comp) * nn + p] = w[3 * p + a];
  }
}
// Dirichlet dofs: unit diagonal (their rows and columns are zero after the probing)
__global__ void k_stencil_dirichlet(int64_t nn, const uint8_t *__restrict__ mask, double *__restrict__ A) {
  for (int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; d < 3 * nn; d += (int64_t)gridDim.x * blockDim.x) {
    if (!mask[d]) continue;
    const int64_t p = d / 3;
    const int a = (int)(d - 3 * p);
    A[((int64_t)(SC * 3 + a) * 3 + a) * nn + p] = 1.0;
  }
}
__global__ void k_stencil_diag(int64_t nn, const double *__restrict__ A, double *__restrict__ d) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * nn; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / 3;
    const int a = (int)(i - 3 * p);
    d[i] = A[((int64_t)(SC * 3 + a) * 3 + a) * nn + p];
  }
}
// y = A x (MatMult semantics: the Dirichlet rows are unit rows), one thread per node, the 125
// neighbours in a fixed order
template <typename T>
__global__ void k_stencil_apply(int NX, int NY, int NZ, const T *__restrict__ A, const T *__restrict__ x,
                                T *__restrict__ y) {
  const int64_t nn = (int64_t)NX * NY * NZ;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nn; p += (int64_t)gridDim.x * blockDim.x) {
    const int I = (int)(p % NX), J = (int)((p / NX) % NY), K = (int)(p / ((int64_t)NX * NY));
    T a0 = T(0), a1 = T(0), a2 = T(0);
    for (int dz = -2; dz <= 2; dz++) {
      if (K + dz < 0 || K + dz >= NZ) continue;
      for (int dy = -2; dy <= 2; dy++) {
        if (J + dy < 0 || J + dy >= NY) continue;
#pragma unroll
        for (int dx = -2; dx <= 2; dx++) {
          if (I + dx < 0 || I + dx >= NX) continue;
          const int64_t q = p + dx + (int64_t)NX * (dy + (int64_t)NY * dz);
          const int o = (dz + 2) * 25 + (dy + 2) * 5 + (dx + 2);
          const T *Ao = A + (int64_t)o * 9 * nn + p;
          const T x0 = x[3 * q], x1 = x[3 * q + 1], x2 = x[3 * q + 2];
          a0 = fma(Ao[0], x0, fma(Ao[nn], x1, fma(Ao[2 * nn], x2, a0)));
          a1 = fma(Ao[3 * nn], x0, fma(Ao[4 * nn], x1, fma(Ao[5 * nn], x2, a1)));
          a2 = fma(Ao[6 * nn], x0, fma(Ao[7 * nn], x1, fma(Ao[8 * nn], x2, a2)));
        }
      }
    }
    y[3 * p] = a0;
    y[3 * p + 1] = a1;
    y[3 * p + 2] = a2;
  }
}
__global__ void k_round_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

// launch shape of a level's smoother apply, picked at set-up by timing (autotune)
struct LevelTune {
  int split = 0;              // kernel: 0 one wave per brick, 1 row-split (3 waves per brick),
                              // 2 matrix-core form (FP32 smoothing only)
  int fast = 0;               // tile addressing of full tiles: 1 = 32-bit fast path, 0 = general
  double ms = 0.0, ms_default = 0.0;  // measured time per apply: chosen shape, default shape
};

struct ptq2_mg {
  ptq2_mg_opts opt{};
  std::vector<LevelTune> tune;          // per level (smoothing levels)
  std::vector<char> gal;                // per level: Galerkin stencil operator (galerkin_levels)
  std::vector<double *> gst64;          // its stencil in FP64 (probing, FP64 applies)
  std::vector<float *> gst32;           // and in FP32 (FP32 smoothing)
  std::vector<void *> gmf_in, gmf_out;  // Galerkin level applied as R A P: vectors of the level above
  std::vector<ptq2::Box> lbox;          // per level: the box of the level's launches (tuned fast32)
  bool prof = false;                    // record per-level events in the eager V-cycle
  std::vector<hipEvent_t> pev;          // 4 per level: entry, before / after the coarser level, exit
  int nl = 0;
  bool f32 = true;
  hipStream_t stream = nullptr;
  std::vector<ptq2_ctx *> ctx;          // level contexts (ctx[0] is the caller's)
  std::vector<bool> owned;              // whether ctx[l] was created here
  std::vector<int> m;                   // (mx,my,mz) triples
  std::vector<int64_t> n;               // velocity dofs per level
  std::vector<void *> vb, vx, vy, vp1, vp2, vdinv;  // level vectors in the smoothing precision
  std::vector<double> emax, emin;
  double *ainv = nullptr;               // dense inverse of the coarsest operator
  int ainv_ld = 0;                      // its leading dimension (padded to the sweep block)
  VcIO *io = nullptr;                   // input/output record of the V-cycle (level 0, FP64 vectors)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // Krylov workspace.  FP64 basis: Vk[j] holds candidate j, then basis vector j.  FP32 basis
  // (basis32): candidates alternate between Vk[0] and Vk[1], finished vectors go to V32[j].
  bool basis32 = false;
  std::vector<double *> Vk, Zk;
  std::vector<float *> V32;
  double *w = nullptr, *part = nullptr, *psum = nullptr, *hdev = nullptr;
  const double **Vptr = nullptr;  // device table of the finished basis vectors (Vk or V32)
  const double **Zptr = nullptr;
  std::vector<double> hist;
  bool stagnated = false;  // the last solve stopped at the rounding floor
  std::string last;
  void *fine_stream0 = nullptr;  // the fine context's stream before the hierarchy was built
  uint64_t fine_gen = 0;         // the fine context's mesh/viscosity generation at set-up
  // rigid-body deflation of stiff inclusions (rg_nc == 0: off or none found)
  int rg_nc = 0, rg_n6 = 0, rg_kept = 0, rg_nchunk = 0;
  int64_t rg_nd = 0;
  std::vector<int> rg_label;  // inclusion of every fine node (-1: none)
  int *rg_nodes = nullptr, *rg_cchunk = nullptr, *rg_dnode = nullptr, *rg_dptr = nullptr, *rg_ecomp = nullptr;
  RgChunk *rg_chunk = nullptr;
  double4 *rg_cent = nullptr;
  double *rg_part = nullptr, *rg_m = nullptr, *rg_y = nullptr, *rg_einv = nullptr, *rg_aw = nullptr;
};

namespace {

int mgfail(ptq2_mg *g, int code, const char *msg) {
  if (g) g->last = msg;
  return code;
}

#define MGCHK(call)                                                                    \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) return mgfail(mg, PTQ2_ERR_DEVICE, hipGetErrorString(e_));   \
  } while (0)
#define MGRC(call)             \
  do {                         \
    int r_ = (call);           \
    if (r_ != PTQ2_OK) return r_; \
  } while (0)

int mg_alloc(ptq2_mg *mg, void **p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return mgfail(mg, PTQ2_ERR_ALLOC, "hipMalloc failed in the multigrid set-up");
  }
  return PTQ2_OK;
}

// y = A x on level l in precision T (MatMult semantics: masked columns, y_D = x_D)
template <typename T>
int level_matmult(ptq2_mg *mg, int l, const T *x, T *y);
// the level's A11 apply in its tuned launch shape (gather: with the face pass)
template <typename T>
hipError_t galerkin_mf(ptq2_mg *mg, int l, const T *x, T *y);

// (the row-split choice is made for the smoothing precision and applies to it only)
hipError_t level_launch(ptq2_mg *mg, int l, const double *x, double *y, bool gather) {
  ptq2_ctx *c = mg->ctx[l];
  const ptq2::Box &b = mg->lbox[l];
  if (mg->gal[l] && mg->tune[l].split == 4 && !mg->f32) return galerkin_mf<double>(mg, l, x, y);
  if (mg->gal[l]) {
    hipLaunchKernelGGL(k_stencil_apply<double>, dim3(nblk(c->nnodes)), dim3(TPB), 0, mg->stream, b.NX, b.NY, b.NZ,
                       (const double *)mg->gst64[l], x, y);
    return hipGetLastError();
  }
  // geometry recomputed from the coordinates (the faster FP64 kernel); the level's face
  // buffer was allocated by its set-up applies (no allocation here: this runs under graph
  // capture)
  if (mg->tune[l].split == 1 && !mg->f32)
    return ptq2::launch_a11_fp64_split(b, c->coords, x, c->mask, c->eta_bm, y, false, mg->stream,
                                       (double *)c->facebuf, gather);
  return ptq2::launch_a11_fp64(b, c->coords, x, c->mask, c->eta_bm, y, false, mg->stream, (double *)c->facebuf,
                               gather);
}
hipError_t level_launch(ptq2_mg *mg, int l, const float *x, float *y, bool gather) {
  ptq2_ctx *c = mg->ctx[l];
  const ptq2::Box &b = mg->lbox[l];
  if (mg->gal[l] && mg->tune[l].split == 4 && mg->f32) return galerkin_mf<float>(mg, l, x, y);
  if (mg->gal[l]) {
    hipLaunchKernelGGL(k_stencil_apply<float>, dim3(nblk(c->nnodes)), dim3(TPB), 0, mg->stream, b.NX, b.NY, b.NZ,
                       (const float *)mg->gst32[l], x, y);
    return hipGetLastError();
  }
  if (mg->tune[l].split == 1 && mg->f32)
    return ptq2::launch_a11_fp32_split(b, x, c->mask, c->fac32, y, false, mg->stream, (float *)c->facebuf, gather);
  if (mg->tune[l].split == 2 && mg->f32)
    return ptq2::launch_a11_fp32_mfma(b, x, c->mask, c->fac32, y, false, mg->stream, (float *)c->facebuf, gather);
  return ptq2::launch_a11_fp32(b, x, c->mask, c->fac32, y, false, mg->stream, (float *)c->facebuf, gather);
}

// A Galerkin level applied matrix-free: y = Pi R A P Pi x + (I - Pi) x with A the (matrix-free)
// operator of the level above, in the smoothing precision, through two vectors of that level
template <typename T>
hipError_t galerkin_mf(ptq2_mg *mg, int l, const T *x, T *y) {
  ptq2_ctx *c = mg->ctx[l], *f = mg->ctx[l - 1];
  T *gp = (T *)mg->gmf_in[l], *ga = (T *)mg->gmf_out[l];
  const int64_t nf = f->nvel;
  const int *mc = &mg->m[3 * l], *mf = &mg->m[3 * (l - 1)];
  const int NXc = 2 * mc[0] + 1, NYc = 2 * mc[1] + 1, NZc = 2 * mc[2] + 1;
  const int NXf = 2 * mf[0] + 1, NYf = 2 * mf[1] + 1, NZf = 2 * mf[2] + 1;
  hipLaunchKernelGGL(k_fill<T>, dim3(nblk(nf)), dim3(TPB), 0, mg->stream, nf, gp, T(0));
  hipLaunchKernelGGL(k_prolong_add<T>, dim3(nblk(nf)), dim3(TPB), 0, mg->stream, NXc, NYc, NXf, NYf, NZf, x,
                     (const uint8_t *)f->mask, gp);
  hipError_t e = level_launch(mg, l - 1, (const T *)gp, ga, true);
  if (e == hipSuccess) e = ptq2::launch_bc_rows<T>(f->ndir, f->dir_idx, (const T *)gp, ga, 0, mg->stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_restrict<T>, dim3(nblk(c->nvel)), dim3(TPB), 0, mg->stream, NXc, NYc, NZc, NXf, NYf, NZf,
                     (const T *)ga, (const T *)nullptr, (const uint8_t *)c->mask, y);
  e = hipGetLastError();
  return e == hipSuccess ? ptq2::launch_bc_rows<T>(c->ndir, c->dir_idx, x, y, 0, mg->stream) : e;
}

template <>
int level_matmult<double>(ptq2_mg *mg, int l, const double *x, double *y) {
  ptq2_ctx *c = mg->ctx[l];
  MGCHK(level_launch(mg, l, x, y, true));
  MGCHK(ptq2::launch_bc_rows<double>(c->ndir, c->dir_idx, x, y, 0, mg->stream));
  return PTQ2_OK;
}
template <>
int level_matmult<float>(ptq2_mg *mg, int l, const float *x, float *y) {
  ptq2_ctx *c = mg->ctx[l];
  MGCHK(level_launch(mg, l, x, y, true));
  MGCHK(ptq2::launch_bc_rows<float>(c->ndir, c->dir_idx, x, y, 0, mg->stream));
  return PTQ2_OK;
}

// the level's A11 apply without its face pass and Dirichlet rows: k_cheb_faces completes y
template <typename T>
int level_apply_owned(ptq2_mg *mg, int l, const T *x, T *y) {
  MGCHK(level_launch(mg, l, x, y, false));
  return PTQ2_OK;
}

// the apply with the Chebyshev update of co fused into its store (owned nodes without face
// contributions); the rest is left in y for k_cheb_faces(face_only)
template <typename T>
int level_apply_cheb(ptq2_mg *mg, int l, const T *x, T *y, const ptq2::ChebOut<T> &co);
template <>
int level_apply_cheb<double>(ptq2_mg *mg, int l, const double *x, double *y, const ptq2::ChebOut<double> &co) {
  ptq2_ctx *c = mg->ctx[l];
  MGCHK(ptq2::launch_a11_fp64(mg->lbox[l], c->coords, x, c->mask, c->eta_bm, y, false, mg->stream, (double *)c->facebuf,
                              false, &co));
  return PTQ2_OK;
}
template <>
int level_apply_cheb<float>(ptq2_mg *mg, int l, const float *x, float *y, const ptq2::ChebOut<float> &co) {
  ptq2_ctx *c = mg->ctx[l];
  MGCHK(ptq2::launch_a11_fp32(mg->lbox[l], x, c->mask, c->fac32, y, false, mg->stream, (float *)c->facebuf, false,
                              &co));
  return PTQ2_OK;
}

// ChebychevRN on level l: x <- smoothed solution of A x = b; zero_guess skips the first residual
template <typename T>
int chebyshev(ptq2_mg *mg, int l, bool zero_guess) {
  const int64_t n = mg->n[l];
  T *b = (T *)mg->vb[l], *x = (T *)mg->vx[l], *y = (T *)mg->vy[l], *dinv = (T *)mg->vdinv[l];
  T *buf[3] = {x, (T *)mg->vp1[l], (T *)mg->vp2[l]};
  const double emax = mg->emax[l], emin = mg->emin[l];
  const double scale = 2.0 / (emax + emin);
  const double alpha = 1.0 - scale * emin;
  const double mu = 1.0 / alpha, omegaprod = 2.0 / alpha;
  double c[3];
  int km1 = 0, k = 1, kp1 = 2;
  c[km1] = 1.0;
  c[k] = mu;
  // apply (owner values + face slots) and update in one pass over the vectors
  const ptq2_ctx *lc = mg->ctx[l];
  // one Chebyshev step from xin: the apply, then the update.  Fused: the apply's store
  // updates the final owned dofs and the face pass the lower-face ones; otherwise the apply
  // stores A xin and the face pass updates every dof.
  const bool fz = mg->opt.fused_smoother != 0 && lc->facebuf && mg->tune[l].split == 0 && !mg->gal[l];
  const T *face = mg->gal[l] ? nullptr : (const T *)lc->facebuf;  // a stencil apply is complete
  auto step = [&](const T *xin, const T *pkm1, T *out, double cr, double ck, double cx) -> int {
    const ptq2::ChebOut<T> co{(const T *)b, (const T *)dinv, xin, pkm1, out, (const uint8_t *)lc->mask, (T)cr, (T)ck,
                              (T)cx};
    if (fz)
      MGRC(level_apply_cheb<T>(mg, l, xin, y, co));
    else
      MGRC(level_apply_owned<T>(mg, l, xin, y));
    hipLaunchKernelGGL(k_cheb_faces<T>, dim3(nblk(n)), dim3(TPB), 0, mg->stream, lc->box, n, (const T *)b,
                       (const T *)y, (const T *)dinv, xin, pkm1, out, face,
                       (const uint8_t *)lc->mask, (T)cr, (T)ck, (T)cx, fz ? 1 : 0);
    return PTQ2_OK;
  };
  if (zero_guess) {
    hipLaunchKernelGGL(k_cheb_first<T>, dim3(nblk(n)), dim3(TPB), 0, mg->stream, n, (const T *)b, (const T *)nullptr,
                       (const T *)dinv, (const T *)nullptr, buf[k], (T)scale);
    hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n)), dim3(TPB), 0, mg->stream, n, x, T(0));
  } else {
    MGRC(step(x, nullptr, buf[k], scale, 0.0, 1.0));
  }
  for (int i = 0; i < mg->opt.smooth_its; i++) {
    c[kp1] = 2.0 * mu * c[k] - c[km1];
    const double omega = omegaprod * c[k] / c[kp1];
    MGRC(step(buf[k], buf[km1], buf[kp1], omega * scale, 1.0 - omega, omega));
    int t = km1;
    km1 = k;
    k = kp1;
    kp1 = t;
  }
  if (k != 0) MGCHK(hipMemcpyAsync(x, buf[k], sizeof(T) * n, hipMemcpyDeviceToDevice, mg->stream));
  return hipGetLastError() == hipSuccess ? PTQ2_OK : mgfail(mg, PTQ2_ERR_DEVICE, "chebyshev launch");
}

// per-level timing (eager V-cycles only): event k of level l
void prof_mark(ptq2_mg *mg, int l, int k) {
  if (mg->prof) (void)hipEventRecord(mg->pev[4 * l + k], mg->stream);
}

template <typename T>
int vcycle_level(ptq2_mg *mg, int l) {
  const int64_t n = mg->n[l];
  T *b = (T *)mg->vb[l], *x = (T *)mg->vx[l], *y = (T *)mg->vy[l];
  prof_mark(mg, l, 0);
  if (l == mg->nl - 1) {
    int nn = (int)n;
    hipLaunchKernelGGL(k_dense_gemv<T>, dim3((nn + 3) / 4), dim3(256), 0, mg->stream, nn, mg->ainv_ld, (const double *)mg->ainv,
                       (const T *)b, x);
    prof_mark(mg, l, 3);
    return hipGetLastError() == hipSuccess ? PTQ2_OK : mgfail(mg, PTQ2_ERR_DEVICE, "coarse solve launch");
  }
  MGRC(chebyshev<T>(mg, l, true));
  MGRC(level_matmult<T>(mg, l, x, y));
  const int *mf = &mg->m[3 * l], *mc = &mg->m[3 * (l + 1)];
  const int NXf = 2 * mf[0] + 1, NYf = 2 * mf[1] + 1, NZf = 2 * mf[2] + 1;
  const int NXc = 2 * mc[0] + 1, NYc = 2 * mc[1] + 1, NZc = 2 * mc[2] + 1;
  hipLaunchKernelGGL(k_restrict<T>, dim3(nblk(mg->n[l + 1])), dim3(TPB), 0, mg->stream, NXc, NYc, NZc, NXf, NYf, NZf,
                     (const T *)b, (const T *)y, (const uint8_t *)mg->ctx[l + 1]->mask, (T *)mg->vb[l + 1]);
  prof_mark(mg, l, 1);
  MGRC(vcycle_level<T>(mg, l + 1));
  prof_mark(mg, l, 2);
  hipLaunchKernelGGL(k_prolong_add<T>, dim3(nblk(n)), dim3(TPB), 0, mg->stream, NXc, NYc, NXf, NYf, NZf,
                     (const T *)mg->vx[l + 1], (const uint8_t *)mg->ctx[l]->mask, x);
  MGRC(chebyshev<T>(mg, l, false));
  prof_mark(mg, l, 3);
  return hipGetLastError() == hipSuccess ? PTQ2_OK : mgfail(mg, PTQ2_ERR_DEVICE, "v-cycle launch");
}

// The Galerkin operator Pi R A P Pi + (I - Pi) of level l through level l - 1 (FP64 MatMult of
// level l-1: its matrix-free operator or its own stencil), assembled by coloured probing: nodes
// whose indices agree modulo 5 on every axis are more than 4 apart, so they couple to disjoint
// stencil rows, and one probe per class and component (375) gives every entry.
int build_galerkin_level(ptq2_mg *mg, int l) {
  ptq2_ctx *c = mg->ctx[l], *f = mg->ctx[l - 1];
  const int64_t nn = c->nnodes, nc = c->nvel, nf = f->nvel;
  const size_t ne = (size_t)nn * SN * 9;
  MGRC(mg_alloc(mg, (void **)&mg->gst64[l], sizeof(double) * ne));
  if (mg->f32) MGRC(mg_alloc(mg, (void **)&mg->gst32[l], sizeof(float) * ne));
  double *e = nullptr, *w = nullptr, *vf = nullptr, *wf = nullptr;
  int rc = mg_alloc(mg, (void **)&e, sizeof(double) * nc);
  if (!rc) rc = mg_alloc(mg, (void **)&w, sizeof(double) * nc);
  if (!rc) rc = mg_alloc(mg, (void **)&vf, sizeof(double) * nf);
  if (!rc) rc = mg_alloc(mg, (void **)&wf, sizeof(double) * nf);
  const int *mc = &mg->m[3 * l], *mf = &mg->m[3 * (l - 1)];
  const int NXc = 2 * mc[0] + 1, NYc = 2 * mc[1] + 1, NZc = 2 * mc[2] + 1;
  const int NXf = 2 * mf[0] + 1, NYf = 2 * mf[1] + 1, NZf = 2 * mf[2] + 1;
  if (!rc) hipLaunchKernelGGL(k_fill<double>, dim3(nblk((int64_t)ne)), dim3(TPB), 0, mg->stream, (int64_t)ne,
                              mg->gst64[l], 0.0);
  int probes = 0;
  for (int cz = 0; cz < 5 && !rc; cz++)
    for (int cy = 0; cy < 5 && !rc; cy++)
      for (int cx = 0; cx < 5 && !rc; cx++)
        for (int comp = 0; comp < 3 && !rc; comp++) {
          hipLaunchKernelGGL(k_probe_vec, dim3(nblk(nc)), dim3(TPB), 0, mg->stream, NXc, NYc, NZc, 5, cx, cy, cz,
                             comp, (const uint8_t *)c->mask, e);
          hipLaunchKernelGGL(k_fill<double>, dim3(nblk(nf)), dim3(TPB), 0, mg->stream, nf, vf, 0.0);
          hipLaunchKernelGGL(k_prolong_add<double>, dim3(nblk(nf)), dim3(TPB), 0, mg->stream, NXc, NYc, NXf, NYf,
                             NZf, (const double *)e, (const uint8_t *)f->mask, vf);
          rc = level_matmult<double>(mg, l - 1, vf, wf);
          hipLaunchKernelGGL(k_restrict<double>, dim3(nblk(nc)), dim3(TPB), 0, mg->stream, NXc, NYc, NZc, NXf, NYf,
                             NZf, (const double *)wf, (const double *)nullptr, (const uint8_t *)c->mask, w);
          hipLaunchKernelGGL(k_stencil_scatter, dim3(nblk(nn)), dim3(TPB), 0, mg->stream, NXc, NYc, NZc, cx, cy, cz,
                             comp, (const uint8_t *)c->mask, (const double *)w, mg->gst64[l]);
          if (!rc && hipGetLastError() != hipSuccess) rc = mgfail(mg, PTQ2_ERR_DEVICE, "galerkin probe launch");
          // bound the work queued ahead of the host
          if (!rc && ++probes % 32 == 0 && hipStreamSynchronize(mg->stream) != hipSuccess)
            rc = mgfail(mg, PTQ2_ERR_DEVICE, "galerkin probing");
        }
  if (!rc && c->mask)
    hipLaunchKernelGGL(k_stencil_dirichlet, dim3(nblk(nc)), dim3(TPB), 0, mg->stream, nn,
                       (const uint8_t *)c->mask, mg->gst64[l]);
  if (!rc && mg->f32)
    hipLaunchKernelGGL(k_round_f32, dim3(nblk((int64_t)ne)), dim3(TPB), 0, mg->stream, (int64_t)ne,
                       (const double *)mg->gst64[l], mg->gst32[l]);
  if (!rc && (hipGetLastError() != hipSuccess || hipStreamSynchronize(mg->stream) != hipSuccess))
    rc = mgfail(mg, PTQ2_ERR_DEVICE, "galerkin level");
  for (double *p : {e, w, vf, wf})
    if (p) (void)hipFree(p);
  if (!rc) mg->gal[l] = 1;
  return rc;
}

// ms per apply of level l's smoother apply (no face pass) in the current launch shape: the
// fastest of 3 timed batches of reps launches, after 2 untimed ones
template <typename T>
int time_level(ptq2_mg *mg, int l, int reps, double *ms) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  MGCHK(hipEventCreate(&e0));
  hipError_t e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return mgfail(mg, PTQ2_ERR_DEVICE, hipGetErrorString(e));
  }
  const T *x = (const T *)mg->vx[l];
  T *y = (T *)mg->vy[l];
  double best = 1e30;
  for (int k = 0; k < 2 && e == hipSuccess; k++) e = level_launch(mg, l, x, y, false);
  for (int k = 0; k < 3 && e == hipSuccess; k++) {
    e = hipEventRecord(e0, mg->stream);
    for (int i = 0; i < reps && e == hipSuccess; i++) e = level_launch(mg, l, x, y, false);
    if (e == hipSuccess) e = hipEventRecord(e1, mg->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e == hipSuccess) best = std::min(best, (double)t / reps);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return mgfail(mg, PTQ2_ERR_DEVICE, hipGetErrorString(e));
  *ms = best;
  return PTQ2_OK;
}

// Autotuner of the smoother apply, per level: times each launch shape (one wave per brick, the
// row-split kernel's three, and for FP32 smoothing the matrix-core form; fast or general tile
// addressing where the tiles are full) and keeps the fastest, the default shape unless another
// is at least 3% faster.  Levels with few bricks (e.g. 16^3: 64 bricks for 1,024 SIMDs) are
// latency-bound in the one-wave kernel.
template <typename T>
int autotune(ptq2_mg *mg) {
  for (int l = 0; l + 1 < mg->nl; l++) {
    ptq2_ctx *c = mg->ctx[l];
    LevelTune &t = mg->tune[l];
    if (mg->gal[l]) {
      // a Galerkin level below a matrix-free one: its stencil, or R A P through the level above
      if (!mg->opt.autotune || mg->gal[l - 1] || mg->opt.autotune == 2 || mg->opt.autotune == 3) continue;
      const size_t vb = sizeof(T) * (size_t)mg->n[l - 1];
      if (!mg->gmf_in[l]) MGRC(mg_alloc(mg, &mg->gmf_in[l], vb));
      if (!mg->gmf_out[l]) MGRC(mg_alloc(mg, &mg->gmf_out[l], vb));
      if (mg->opt.autotune == 4) {
        t.split = 4;
        continue;
      }
      const int64_t n = mg->n[l];
      hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n)), dim3(TPB), 0, mg->stream, n, (T *)mg->vx[l], T(1));
      double t0 = 0.0;
      MGRC(time_level<T>(mg, l, 2, &t0));
      const int reps = (int)std::min(50.0, std::max(2.0, 2.0 / std::max(t0, 1e-3)));
      double ms_st = 0.0, ms_mf = 0.0;
      t.split = 0;
      MGRC(time_level<T>(mg, l, reps, &ms_st));
      t.split = 4;
      MGRC(time_level<T>(mg, l, reps, &ms_mf));
      t.ms_default = ms_st;
      t.split = ms_mf < 0.97 * ms_st ? 4 : 0;
      t.ms = t.split ? ms_mf : ms_st;
      continue;
    }
    if (!mg->opt.autotune || !c->facebuf) continue;
    if (mg->opt.autotune == 4) continue;  // the one-wave kernel on the matrix-free levels
    if (mg->opt.autotune >= 2) {  // one kernel everywhere, untimed
      t.split = mg->opt.autotune == 2 ? 1 : (sizeof(T) == 4 ? 2 : 0);
      continue;
    }
    const int64_t n = mg->n[l];
    hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n)), dim3(TPB), 0, mg->stream, n, (T *)mg->vx[l], T(1));
    double t0 = 0.0;
    MGRC(time_level<T>(mg, l, 2, &t0));
    const int reps = (int)std::min(50.0, std::max(2.0, 2.0 / std::max(t0, 1e-3)));  // ~2 ms per batch
    double best = 1e30;
    int bs = 0, bf = c->box.fast32;
    const int kinds = sizeof(T) == 4 ? 3 : 2;
    for (int split = 0; split < kinds; split++)
      for (int fast = c->box.fast32; fast >= 0; fast--) {
        t.split = split;
        mg->lbox[l].fast32 = fast;
        double ms = 0.0;
        MGRC(time_level<T>(mg, l, reps, &ms));
        const bool dflt = split == 0 && fast == c->box.fast32;
        if (dflt) t.ms_default = ms;
        if (dflt ? ms < best / 0.97 : ms < 0.97 * best) {
          best = ms;
          bs = split;
          bf = fast;
        }
      }
    t.split = bs;
    t.fast = bf;
    t.ms = best;
    mg->lbox[l].fast32 = bf;
  }
  return PTQ2_OK;
}

// A-DEF2 steps around the V-cycle on the level-0 vectors: before it b -= A W y with
// y = E^-1 W^T b, after it x += W y
template <typename T>
void rigid_pre(ptq2_mg *mg, T *b) {
  if (!mg->rg_nc) return;
  const ptq2_ctx *f = mg->ctx[0];
  hipLaunchKernelGGL(k_rg_moments<T>, dim3(mg->rg_nchunk), dim3(TPB), 0, mg->stream, (const T *)b,
                     (const uint8_t *)f->mask, (const double *)f->coords, (const int *)mg->rg_nodes,
                     (const RgChunk *)mg->rg_chunk, (const double4 *)mg->rg_cent, mg->rg_part);
  hipLaunchKernelGGL(k_rg_msum, dim3(mg->rg_nc), dim3(TPB), 0, mg->stream, (const int *)mg->rg_cchunk,
                     (const double *)mg->rg_part, mg->rg_m);
  hipLaunchKernelGGL(k_rg_coarse, dim3((mg->rg_n6 + 3) / 4), dim3(256), 0, mg->stream, mg->rg_n6,
                     (const double *)mg->rg_m, (const double *)mg->rg_einv, mg->rg_y);
  hipLaunchKernelGGL(k_rg_subtract<T>, dim3(nblk(mg->rg_nd)), dim3(TPB), 0, mg->stream, mg->rg_nd,
                     (const int *)mg->rg_dnode, (const int *)mg->rg_dptr, (const int *)mg->rg_ecomp,
                     (const double *)mg->rg_aw, (const double *)mg->rg_y, b);
}
template <typename T>
void rigid_post(ptq2_mg *mg, T *x) {
  if (!mg->rg_nc) return;
  const ptq2_ctx *f = mg->ctx[0];
  hipLaunchKernelGGL(k_rg_expand<T>, dim3(mg->rg_nchunk), dim3(TPB), 0, mg->stream, (const uint8_t *)f->mask,
                     (const double *)f->coords, (const int *)mg->rg_nodes, (const RgChunk *)mg->rg_chunk,
                     (const double4 *)mg->rg_cent, (const double *)mg->rg_y, x);
}

// the work of one V-cycle: io->out = io->sout * M^-1 (io->sin * io->in), FP64 vectors
int vcycle_body(ptq2_mg *mg) {
  const int64_t n = mg->n[0];
  if (mg->f32) {
    hipLaunchKernelGGL(k_vin<float>, dim3(nblk(n)), dim3(TPB), 0, mg->stream, (const VcIO *)mg->io, n,
                       (float *)mg->vb[0]);
    rigid_pre<float>(mg, (float *)mg->vb[0]);
    MGRC(vcycle_level<float>(mg, 0));
    rigid_post<float>(mg, (float *)mg->vx[0]);
    hipLaunchKernelGGL(k_vout<float>, dim3(nblk(n)), dim3(TPB), 0, mg->stream, (const VcIO *)mg->io, n,
                       (const float *)mg->vx[0]);
  } else {
    hipLaunchKernelGGL(k_vin<double>, dim3(nblk(n)), dim3(TPB), 0, mg->stream, (const VcIO *)mg->io, n,
                       (double *)mg->vb[0]);
    rigid_pre<double>(mg, (double *)mg->vb[0]);
    MGRC(vcycle_level<double>(mg, 0));
    rigid_post<double>(mg, (double *)mg->vx[0]);
    hipLaunchKernelGGL(k_vout<double>, dim3(nblk(n)), dim3(TPB), 0, mg->stream, (const VcIO *)mg->io, n,
                       (const double *)mg->vx[0]);
  }
  return hipGetLastError() == hipSuccess ? PTQ2_OK : mgfail(mg, PTQ2_ERR_DEVICE, "v-cycle");
}

// Reproducible dot products out[v] = V[v].w (v < nv) and, with self, out[nv] = w.w, in one
// pass over w per MD vectors; the partials are summed on the device in a fixed order.
int mdot(ptq2_mg *mg, const double *const *Vdev, int nv, const double *w, int64_t n, bool self, double *out) {
  const int np = nv + (self ? 1 : 0);
  if (np == 0) return PTQ2_OK;
  if (mg->basis32 && nv > 0)
    hipLaunchKernelGGL(k_mdot<float>, dim3(MDB), dim3(TPB), 0, mg->stream, n, (const float *const *)Vdev, nv, w,
                       self ? 1 : 0, mg->part);
  else
    hipLaunchKernelGGL(k_mdot<double>, dim3(MDB), dim3(TPB), 0, mg->stream, n, Vdev, nv, w, self ? 1 : 0, mg->part);
  hipLaunchKernelGGL(k_sum_parts, dim3(np), dim3(TPB), 0, mg->stream, (const double *)mg->part, mg->psum);
  MGCHK(hipMemcpyAsync(out, mg->psum, sizeof(double) * np, hipMemcpyDeviceToHost, mg->stream));
  MGCHK(hipStreamSynchronize(mg->stream));
  return PTQ2_OK;
}

// pass 1 of the delayed-reorthogonalisation step: out[2v], out[2v+1] = V[v].u', V[v].w (v < nv),
// out[2nv..2nv+2] = u'.w, u'.u', w.w
int mdot2(ptq2_mg *mg, int nv, const double *u, const double *w, int64_t n, double *out) {
  const int np = 2 * nv + 3;
  if (mg->basis32)
    hipLaunchKernelGGL(k_mdot2_f32, dim3(MDB), dim3(TPB), 0, mg->stream, n, (const float *const *)mg->Vptr, nv, u, w,
                       mg->part);
  else
    hipLaunchKernelGGL(k_mdot2<double>, dim3(MDB), dim3(TPB), 0, mg->stream, n, mg->Vptr, nv, u, w, mg->part);
  hipLaunchKernelGGL(k_sum_parts, dim3(np), dim3(TPB), 0, mg->stream, (const double *)mg->part, mg->psum);
  MGCHK(hipMemcpyAsync(out, mg->psum, sizeof(double) * np, hipMemcpyDeviceToHost, mg->stream));
  MGCHK(hipStreamSynchronize(mg->stream));
  return PTQ2_OK;
}

// pass 2 (coefficients [a, b] already in hdev): the finished vector into vo (the basis
// precision), the next candidate into wo; returns ||wo||
int dcgs_update(ptq2_mg *mg, int nv, double su, double ia, double gam, const double *u, void *vo, double *wo,
                int64_t n, double *nrm) {
  if (mg->basis32)
    hipLaunchKernelGGL(k_dcgs_update_f32, dim3(MDB), dim3(TPB), 0, mg->stream, n, (const float *const *)mg->Vptr, nv,
                       (const double *)mg->hdev, su, ia, gam, u, (float *)vo, (const double *)mg->w, wo, mg->part);
  else
    hipLaunchKernelGGL(k_dcgs_update<double>, dim3(MDB), dim3(TPB), 0, mg->stream, n, mg->Vptr, nv,
                       (const double *)mg->hdev, su, ia, gam, u, (double *)vo, (const double *)mg->w, wo, mg->part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(TPB), 0, mg->stream, (const double *)mg->part, mg->psum);
  double d;
  MGCHK(hipMemcpyAsync(&d, mg->psum, sizeof(double), hipMemcpyDeviceToHost, mg->stream));
  MGCHK(hipStreamSynchronize(mg->stream));
  *nrm = std::sqrt(std::max(d, 0.0));
  return PTQ2_OK;
}

int norm2(ptq2_mg *mg, const double *x, int64_t n, double *out) {
  double d;
  MGRC(mdot(mg, nullptr, 0, x, n, true, &d));
  *out = std::sqrt(std::max(d, 0.0));
  return PTQ2_OK;
}

// t -= sum_{v<nv} hdev[v] V[v] and the 2-norm of the result, in one pass
int maxpy_norm(ptq2_mg *mg, int nv, double *t, int64_t n, double *nrm) {
  if (mg->basis32)
    hipLaunchKernelGGL(k_maxpy_norm<float>, dim3(MDB), dim3(TPB), 0, mg->stream, n, (const float *const *)mg->Vptr, nv,
                       (const double *)mg->hdev, t, mg->part);
  else
    hipLaunchKernelGGL(k_maxpy_norm<double>, dim3(MDB), dim3(TPB), 0, mg->stream, n, mg->Vptr, nv,
                       (const double *)mg->hdev, t, mg->part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(TPB), 0, mg->stream, (const double *)mg->part, mg->psum);
  double d;
  MGCHK(hipMemcpyAsync(&d, mg->psum, sizeof(double), hipMemcpyDeviceToHost, mg->stream));
  MGCHK(hipStreamSynchronize(mg->stream));
  *nrm = std::sqrt(std::max(d, 0.0));
  return PTQ2_OK;
}

// ---------------------------------------------------------------- coarse inverse
// Dense inverse of the coarsest operator by the block sweep (Gauss-Jordan elimination
// without pivoting, valid for SPD matrices): sweeping diagonal block k with D = A_kk^-1
// maps A_kk -> -D, A_ik -> A_ik D, A_kj -> D A_kj, A_ij -> A_ij - A_ik D A_kj; after all
// blocks the matrix holds -A^-1.  Blocks of SB = 64; the matrix is padded to a multiple of
// SB with identity rows.  Every output has one writer and a fixed summation order, so the
// result is deterministic.
constexpr int SB = 64;


// symmetrise the round-off of the assembled operator
__global__ void k_symmetrise(int np, double *__restrict__ A) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)np * np) return;
  int i = (int)(idx / np), j = (int)(idx % np);
  if (i < j) {
    double s = 0.5 * (A[idx] + A[(int64_t)j * np + i]);
    A[idx] = s;
    A[(int64_t)j * np + i] = s;
  }
}
// sweep of the SB x SB diagonal block k in LDS: Dm = -A_kk^-1; flags a non-positive pivot
__global__ __launch_bounds__(256) void k_sweep_diag(int np, int k, const double *__restrict__ A,
                                                    double *__restrict__ Dm, int *__restrict__ err) {
  __shared__ double S[SB][SB + 1];
  __shared__ double piv_s;
  const int t = threadIdx.x;
  const int64_t o = (int64_t)k * SB;
  for (int e = t; e < SB * SB; e += 256) S[e / SB][e % SB] = A[(o + e / SB) * np + o + e % SB];
  __syncthreads();
  for (int p = 0; p < SB; p++) {
    if (t == 0) {
      piv_s = S[p][p];
      if (!(piv_s > 0.0)) atomicOr(err, 1);
    }
    __syncthreads();
    const double piv = piv_s;
    const double rp = piv > 0.0 ? 1.0 / piv : 0.0;
    // a_ij -= a_ip a_pj / a_pp for i,j != p (reads row p and column p, which stay unchanged here)
    for (int e = t; e < SB * SB; e += 256) {
      const int i = e / SB, j = e % SB;
      if (i != p && j != p) S[i][j] -= S[i][p] * S[p][j] * rp;
    }
    __syncthreads();
    for (int e = t; e < SB; e += 256)
      if (e != p) {
        S[e][p] *= rp;
        S[p][e] *= rp;
      }
    if (t == 0) S[p][p] = -rp;
    __syncthreads();
  }
  for (int e = t; e < SB * SB; e += 256) Dm[e] = S[e / SB][e % SB];
}
// panel of block column k: C_i = A_ik (old), P_i = A_ik D = -A_ik Dm for every block row i != k
__global__ __launch_bounds__(256) void k_sweep_panel(int np, int k, const double *__restrict__ A,
                                                     const double *__restrict__ Dm, double *__restrict__ P,
                                                     double *__restrict__ C) {
  __shared__ double Ds[SB][SB + 1];
  __shared__ double As[4][SB];
  const int t = threadIdx.x;
  for (int e = t; e < SB * SB; e += 256) Ds[e / SB][e % SB] = Dm[e];
  const int r0 = blockIdx.x * 4;  // 4 rows per block, 64 threads per row
  const int rl = t / SB, col = t % SB;
  const int64_t row = r0 + rl;
  const int64_t o = (int64_t)k * SB;
  const bool live = row < np && row / SB != k;
  double a = 0.0;
  if (live) a = A[row * np + o + col];
  As[rl][col] = a;
  __syncthreads();
  if (!live) return;
  double s = 0.0;
  for (int q = 0; q < SB; q++) s = fma(As[rl][q], Ds[q][col], s);
  C[row * SB + col] = a;
  P[row * SB + col] = -s;
}
// rank-SB update of every tile outside block row/column k, and the new block row/column k
__global__ __launch_bounds__(256) void k_sweep_update(int np, int k, double *__restrict__ A,
                                                      const double *__restrict__ P, const double *__restrict__ C,
                                                      const double *__restrict__ Dm) {
  constexpr int KC = SB / 2;  // the rank-SB update in two K chunks keeps LDS at 33 KB
  __shared__ double Ps[SB][KC + 1];
  __shared__ double Cs[SB][KC + 1];
  const int bi = blockIdx.y, bj = blockIdx.x;
  const int t = threadIdx.x;
  const int64_t i0 = (int64_t)bi * SB, j0 = (int64_t)bj * SB;
  if (bi == k || bj == k) {
    for (int e = t; e < SB * SB; e += 256) {
      const int a = e / SB, c = e % SB;
      double v;
      if (bi == k && bj == k)
        v = Dm[e];
      else if (bj == k)
        v = P[(i0 + a) * SB + c];  // A_ik' = P_i
      else
        v = P[(j0 + c) * SB + a];  // A_kj' = P_j^T
      A[(i0 + a) * np + j0 + c] = v;
    }
    return;
  }
  // thread owns a 4x4 sub-tile: rows 4*(t/16).., cols 4*(t%16)..
  const int ra = 4 * (t / 16), ca = 4 * (t % 16);
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[a][c] = A[(i0 + ra + a) * np + j0 + ca + c];
  for (int q0 = 0; q0 < SB; q0 += KC) {
    __syncthreads();
    for (int e = t; e < SB * KC; e += 256) {
      Ps[e / KC][e % KC] = P[(i0 + e / KC) * SB + q0 + e % KC];
      Cs[e / KC][e % KC] = C[(j0 + e / KC) * SB + q0 + e % KC];
    }
    __syncthreads();
    for (int q = 0; q < KC; q++) {
      double pa[4], cc[4];
#pragma unroll
      for (int a = 0; a < 4; a++) pa[a] = Ps[ra + a][q];
#pragma unroll
      for (int c = 0; c < 4; c++) cc[c] = Cs[ca + c][q];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[a][c] = fma(-pa[a], cc[c], acc[a][c]);
    }
  }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 4; c++) A[(i0 + ra + a) * np + j0 + ca + c] = acc[a][c];
}
__global__ void k_negate(int64_t n, double *__restrict__ x) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = -x[i];
}


// Rigid-body deflation set-up on the fine level: find the inclusions, build the node lists,
// then E = W^T A W and the A W entries column by column (one fine apply per column), and
// E^-1 over the columns a left-looking Cholesky factorisation keeps (a column whose pivot
// is not above rel * E_jj -- fully constrained or dependent -- gets a zero row and column).
int rigid_setup(ptq2_mg *mg) {
  ptq2_ctx *f = mg->ctx[0];
  const int mx = f->box.mx, my = f->box.my;  // element layers in z: ne / (mx my)
  const int NX = 2 * mx + 1, NY = 2 * my + 1;
  const int64_t ne = f->nel, nn = f->nnodes, n = f->nvel;
  std::vector<double> emx((size_t)ne), emn((size_t)ne);
  {
    double *d = nullptr;
    MGRC(mg_alloc(mg, (void **)&d, sizeof(double) * 2 * (size_t)ne));
    hipLaunchKernelGGL(k_eta_elem_range, dim3(nblk(ne)), dim3(TPB), 0, mg->stream, f->box, (const double *)f->eta_bm, d,
                       d + ne);
    hipError_t e1 = hipMemcpyAsync(emx.data(), d, sizeof(double) * ne, hipMemcpyDeviceToHost, mg->stream);
    hipError_t e2 = hipMemcpyAsync(emn.data(), d + ne, sizeof(double) * ne, hipMemcpyDeviceToHost, mg->stream);
    hipError_t e3 = hipStreamSynchronize(mg->stream);
    (void)hipFree(d);
    MGCHK(e1);
    MGCHK(e2);
    MGCHK(e3);
  }
  const double lo = *std::min_element(emn.begin(), emn.end()), hi = *std::max_element(emx.begin(), emx.end());
  mg->rg_label.assign((size_t)nn, -1);
  if (!(hi >= mg->opt.rigid_min_contrast * lo)) return PTQ2_OK;
  const double tau = std::sqrt(lo * hi);
  // connected stiff elements (26 neighbours) by union-find, numbered by their first element
  std::vector<int64_t> par((size_t)ne, -1);
  for (int64_t e = 0; e < ne; e++)
    if (emx[e] > tau) par[e] = e;
  auto find = [&](int64_t a) {
    while (par[a] != a) {
      par[a] = par[par[a]];
      a = par[a];
    }
    return a;
  };
  for (int64_t e = 0; e < ne; e++) {
    if (par[e] < 0) continue;
    const int ex = (int)(e % mx), ey = (int)((e / mx) % my), ez = (int)(e / ((int64_t)mx * my));
    for (int dz = -1; dz <= 0; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          if (dz == 0 && (dy > 0 || (dy == 0 && dx >= 0))) continue;  // the 13 earlier neighbours
          const int x = ex + dx, y = ey + dy, z = ez + dz;
          if (x < 0 || x >= mx || y < 0 || y >= my || z < 0) continue;
          const int64_t g = x + (int64_t)mx * (y + (int64_t)my * z);
          if (par[g] < 0) continue;
          const int64_t ra = find(e), rb = find(g);
          if (ra != rb) par[std::max(ra, rb)] = std::min(ra, rb);
        }
  }
  std::vector<int> lab((size_t)ne, -1), size;
  std::vector<int64_t> rid((size_t)ne, -1);
  for (int64_t e = 0; e < ne; e++) {
    if (par[e] < 0) continue;
    const int64_t r = find(e);
    if (rid[r] < 0) {
      rid[r] = (int64_t)size.size();
      size.push_back(0);
    }
    lab[e] = (int)rid[r];
    size[lab[e]]++;
  }
  int nc = (int)size.size();
  if (nc > mg->opt.rigid_max_inclusions) {
    // keep the largest (ties: the earlier), renumbered in their original order
    std::vector<int> ord(nc);
    for (int c = 0; c < nc; c++) ord[c] = c;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return size[a] > size[b]; });
    std::vector<int> keep(ord.begin(), ord.begin() + mg->opt.rigid_max_inclusions), map(nc, -1);
    std::sort(keep.begin(), keep.end());
    for (size_t k = 0; k < keep.size(); k++) map[keep[k]] = (int)k;
    for (int64_t e = 0; e < ne; e++)
      if (lab[e] >= 0) lab[e] = map[lab[e]];
    nc = mg->opt.rigid_max_inclusions;
  }
  if (nc == 0) return PTQ2_OK;
  auto enode = [&](int64_t e, int q) {
    const int ex = (int)(e % mx), ey = (int)((e / mx) % my), ez = (int)(e / ((int64_t)mx * my));
    return (int64_t)(2 * ex + q % 3) + (int64_t)NX * ((2 * ey + (q / 3) % 3) + (int64_t)NY * (2 * ez + q / 9));
  };
  for (int64_t e = 0; e < ne; e++)
    if (lab[e] >= 0)
      for (int q = 0; q < 27; q++) mg->rg_label[enode(e, q)] = lab[e];
  // nodes of every inclusion in increasing order, and their chunks
  std::vector<int> cptr(nc + 1, 0), nodes;
  for (int64_t i = 0; i < nn; i++)
    if (mg->rg_label[i] >= 0) cptr[mg->rg_label[i] + 1]++;
  for (int c = 0; c < nc; c++) cptr[c + 1] += cptr[c];
  nodes.resize(cptr[nc]);
  {
    std::vector<int> at(cptr.begin(), cptr.end() - 1);
    for (int64_t i = 0; i < nn; i++)
      if (mg->rg_label[i] >= 0) nodes[at[mg->rg_label[i]]++] = (int)i;
  }
  std::vector<RgChunk> chunks;
  std::vector<int> cchunk(nc + 1, 0);
  for (int c = 0; c < nc; c++) {
    for (int b = cptr[c]; b < cptr[c + 1]; b += RG_CH) chunks.push_back({c, b, std::min(b + RG_CH, cptr[c + 1]), 0});
    cchunk[c + 1] = (int)chunks.size();
  }
  // centre and RMS radius from the coordinates
  std::vector<double> X((size_t)n);
  MGCHK(hipMemcpy(X.data(), f->coords, sizeof(double) * n, hipMemcpyDeviceToHost));
  std::vector<double4> cent(nc);
  for (int c = 0; c < nc; c++) {
    double s[3] = {0.0, 0.0, 0.0}, q = 0.0;
    const int cnt = cptr[c + 1] - cptr[c];
    for (int i = cptr[c]; i < cptr[c + 1]; i++)
      for (int a = 0; a < 3; a++) s[a] += X[3 * (int64_t)nodes[i] + a];
    for (int a = 0; a < 3; a++) s[a] /= cnt;
    for (int i = cptr[c]; i < cptr[c + 1]; i++)
      for (int a = 0; a < 3; a++) {
        const double d = X[3 * (int64_t)nodes[i] + a] - s[a];
        q += d * d;
      }
    cent[c] = make_double4(s[0], s[1], s[2], 1.0 / std::sqrt(q / cnt));
  }
  // (node, inclusion) pairs reached by A W: the nodes of every element touching an inclusion
  std::vector<uint64_t> pr;
  for (int64_t e = 0; e < ne; e++) {
    int seen[27], ns = 0;
    for (int q = 0; q < 27; q++) {
      const int c = mg->rg_label[enode(e, q)];
      if (c < 0) continue;
      bool dup = false;
      for (int k = 0; k < ns; k++) dup = dup || seen[k] == c;
      if (!dup) seen[ns++] = c;
    }
    for (int k = 0; k < ns; k++)
      for (int q = 0; q < 27; q++) pr.push_back(((uint64_t)enode(e, q) << 8) | (uint64_t)seen[k]);
  }
  std::sort(pr.begin(), pr.end());
  pr.erase(std::unique(pr.begin(), pr.end()), pr.end());
  std::vector<int> dnode, dptr, ecomp(pr.size());
  for (size_t i = 0; i < pr.size(); i++) {
    const int node = (int)(pr[i] >> 8);
    if (dnode.empty() || dnode.back() != node) {
      dnode.push_back(node);
      dptr.push_back((int)i);
    }
    ecomp[i] = (int)(pr[i] & 0xFF);
  }
  dptr.push_back((int)pr.size());
  const int n6 = 6 * nc;
  mg->rg_nc = nc;
  mg->rg_n6 = n6;
  mg->rg_nchunk = (int)chunks.size();
  mg->rg_nd = (int64_t)dnode.size();
  auto up = [&](void **dst, const void *src, size_t bytes) -> int {
    MGRC(mg_alloc(mg, dst, bytes));
    MGCHK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return PTQ2_OK;
  };
  MGRC(up((void **)&mg->rg_nodes, nodes.data(), sizeof(int) * nodes.size()));
  MGRC(up((void **)&mg->rg_chunk, chunks.data(), sizeof(RgChunk) * chunks.size()));
  MGRC(up((void **)&mg->rg_cchunk, cchunk.data(), sizeof(int) * cchunk.size()));
  MGRC(up((void **)&mg->rg_cent, cent.data(), sizeof(double4) * cent.size()));
  MGRC(up((void **)&mg->rg_dnode, dnode.data(), sizeof(int) * dnode.size()));
  MGRC(up((void **)&mg->rg_dptr, dptr.data(), sizeof(int) * dptr.size()));
  MGRC(up((void **)&mg->rg_ecomp, ecomp.data(), sizeof(int) * ecomp.size()));
  MGRC(mg_alloc(mg, (void **)&mg->rg_part, sizeof(double) * 6 * chunks.size()));
  MGRC(mg_alloc(mg, (void **)&mg->rg_y, sizeof(double) * n6));
  MGRC(mg_alloc(mg, (void **)&mg->rg_m, sizeof(double) * n6));
  MGRC(mg_alloc(mg, (void **)&mg->rg_einv, sizeof(double) * n6 * n6));
  MGRC(mg_alloc(mg, (void **)&mg->rg_aw, sizeof(double) * 18 * pr.size()));
  // columns of E and of A W
  double *w = nullptr, *v = nullptr;
  MGRC(mg_alloc(mg, (void **)&w, sizeof(double) * n));
  int rc = mg_alloc(mg, (void **)&v, sizeof(double) * n);
  for (int j = 0; j < n6 && rc == PTQ2_OK; j++) {
    hipLaunchKernelGGL(k_unit, dim3(nblk(n6)), dim3(TPB), 0, mg->stream, (int64_t)n6, (int64_t)j, mg->rg_y);
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(n)), dim3(TPB), 0, mg->stream, n, w, 0.0);
    rigid_post<double>(mg, w);
    rc = ptq2_a11_apply(f, w, v, PTQ2_DEVICE | PTQ2_MATMULT);
    if (rc) break;
    hipLaunchKernelGGL(k_rg_moments<double>, dim3(mg->rg_nchunk), dim3(TPB), 0, mg->stream, (const double *)v,
                       (const uint8_t *)f->mask, (const double *)f->coords, (const int *)mg->rg_nodes,
                       (const RgChunk *)mg->rg_chunk, (const double4 *)mg->rg_cent, mg->rg_part);
    hipLaunchKernelGGL(k_rg_msum, dim3(nc), dim3(TPB), 0, mg->stream, (const int *)mg->rg_cchunk,
                       (const double *)mg->rg_part, mg->rg_einv + (int64_t)j * n6);
    hipLaunchKernelGGL(k_rg_gather, dim3(nblk(mg->rg_nd)), dim3(TPB), 0, mg->stream, mg->rg_nd,
                       (const int *)mg->rg_dnode, (const int *)mg->rg_dptr, (const int *)mg->rg_ecomp, j / 6, j % 6,
                       (const double *)v, mg->rg_aw);
    if (hipGetLastError() != hipSuccess) rc = mgfail(mg, PTQ2_ERR_DEVICE, "rigid deflation set-up launch");
  }
  std::vector<double> E((size_t)n6 * n6);
  if (rc == PTQ2_OK &&
      hipMemcpyAsync(E.data(), mg->rg_einv, sizeof(double) * E.size(), hipMemcpyDeviceToHost, mg->stream) != hipSuccess)
    rc = mgfail(mg, PTQ2_ERR_DEVICE, "rigid deflation copy");
  if (hipStreamSynchronize(mg->stream) != hipSuccess && rc == PTQ2_OK) rc = mgfail(mg, PTQ2_ERR_DEVICE, "rigid deflation");
  (void)hipFree(w);
  if (v) (void)hipFree(v);
  MGRC(rc);
  for (int i = 0; i < n6; i++)
    for (int j = 0; j < i; j++) E[(size_t)i * n6 + j] = E[(size_t)j * n6 + i] = 0.5 * (E[(size_t)i * n6 + j] + E[(size_t)j * n6 + i]);
  // left-looking Cholesky over the kept columns
  const double rel = 1e-12;
  std::vector<int> keep;
  std::vector<double> L((size_t)n6 * n6, 0.0), l(n6);
  for (int j = 0; j < n6; j++) {
    const double ejj = E[(size_t)j * n6 + j];
    if (!(ejj > 0.0)) continue;
    const int k = (int)keep.size();
    double piv = ejj;
    for (int i = 0; i < k; i++) {
      double t = E[(size_t)keep[i] * n6 + j];
      for (int q = 0; q < i; q++) t -= L[(size_t)i * n6 + q] * l[q];
      l[i] = t / L[(size_t)i * n6 + i];
      piv -= l[i] * l[i];
    }
    if (!(piv > rel * ejj)) continue;
    for (int q = 0; q < k; q++) L[(size_t)k * n6 + q] = l[q];
    L[(size_t)k * n6 + k] = std::sqrt(piv);
    keep.push_back(j);
  }
  const int nk = (int)keep.size();
  // E_kept^-1 = L^-T L^-1, scattered into the full index range
  std::vector<double> Li((size_t)nk * nk, 0.0), Ei((size_t)n6 * n6, 0.0);
  for (int c = 0; c < nk; c++) {
    Li[(size_t)c * nk + c] = 1.0 / L[(size_t)c * n6 + c];
    for (int r = c + 1; r < nk; r++) {
      double t = 0.0;
      for (int q = c; q < r; q++) t += L[(size_t)r * n6 + q] * Li[(size_t)q * nk + c];
      Li[(size_t)r * nk + c] = -t / L[(size_t)r * n6 + r];
    }
  }
  for (int a = 0; a < nk; a++)
    for (int b = 0; b <= a; b++) {
      double t = 0.0;
      for (int q = a; q < nk; q++) t += Li[(size_t)q * nk + a] * Li[(size_t)q * nk + b];
      Ei[(size_t)keep[a] * n6 + keep[b]] = Ei[(size_t)keep[b] * n6 + keep[a]] = t;
    }
  mg->rg_kept = nk;
  MGCHK(hipMemcpy(mg->rg_einv, Ei.data(), sizeof(double) * Ei.size(), hipMemcpyHostToDevice));
  return PTQ2_OK;
}


// z = M^-1 (scale r): one V-cycle reading r and writing z directly (the input is read in full
// by the graph's first kernel before its last kernel writes z, so r and z may alias)
int vcycle_scaled(ptq2_mg *mg, const double *r, double scale, double *z) {
  hipLaunchKernelGGL(k_set_io, dim3(1), dim3(64), 0, mg->stream, mg->io, r, z, scale, 1.0);
  if (mg->exec) MGCHK(hipGraphLaunch(mg->exec, mg->stream));
  else MGRC(vcycle_body(mg));
  return hipGetLastError() == hipSuccess ? PTQ2_OK : mgfail(mg, PTQ2_ERR_DEVICE, "v-cycle launch");
}

}  // namespace

extern "C" {

void ptq2_mg_default_opts(ptq2_mg_opts *o) {
  if (!o) return;
  o->nlevels = 5;
  o->smooth_its = 4;
  o->fp32_smoother = 1;
  o->emin_frac = 0.2;
  o->emax_frac = 1.1;
  o->esteig_its = 20;
  o->use_graph = 1;
  o->coarse_max_dofs = 20000;
  o->galerkin_coarse = 1;
  o->coarse_eta_mean = PTQ2_ETA_MEAN_ARITHMETIC;
  o->fp32_basis = 0;
  o->rigid_deflation = 0;
  o->rigid_min_contrast = 100.0;
  o->rigid_max_inclusions = RG_MAXINC;
  o->fused_smoother = 0;
  o->autotune = 1;
  o->galerkin_levels = 0;
}

int ptq2_mg_destroy(ptq2_mg **pmg) {
  if (!pmg || !*pmg) return PTQ2_OK;
  ptq2_mg *mg = *pmg;
  if (mg->stream) (void)hipStreamSynchronize(mg->stream);
  for (hipEvent_t e : mg->pev)
    if (e) (void)hipEventDestroy(e);
  for (double *p : mg->gst64)
    if (p) (void)hipFree(p);
  for (float *p : mg->gst32)
    if (p) (void)hipFree(p);
  for (void *p : mg->gmf_in)
    if (p) (void)hipFree(p);
  for (void *p : mg->gmf_out)
    if (p) (void)hipFree(p);
  if (mg->exec) (void)hipGraphExecDestroy(mg->exec);
  if (mg->graph) (void)hipGraphDestroy(mg->graph);
  for (size_t l = 0; l < mg->vb.size(); l++) {
    for (void *p : {mg->vb[l], mg->vx[l], mg->vy[l], mg->vp1[l], mg->vp2[l], mg->vdinv[l]})
      if (p) (void)hipFree(p);
  }
  for (size_t l = 1; l < mg->ctx.size(); l++)
    if (mg->owned[l]) ptq2_destroy(&mg->ctx[l]);
  if (!mg->ctx.empty() && mg->ctx[0]) ptq2_set_stream(mg->ctx[0], mg->fine_stream0);
  for (double *p : mg->Vk) (void)hipFree(p);
  for (float *p : mg->V32) (void)hipFree(p);
  for (double *p : mg->Zk) (void)hipFree(p);
  for (void *p : {(void *)mg->ainv, (void *)mg->io, (void *)mg->w, (void *)mg->part, (void *)mg->psum,
                  (void *)mg->hdev, (void *)mg->Vptr, (void *)mg->Zptr, (void *)mg->rg_nodes, (void *)mg->rg_cchunk,
                  (void *)mg->rg_dnode, (void *)mg->rg_dptr, (void *)mg->rg_ecomp, (void *)mg->rg_chunk,
                  (void *)mg->rg_cent, (void *)mg->rg_part, (void *)mg->rg_m, (void *)mg->rg_y, (void *)mg->rg_einv, (void *)mg->rg_aw})
    if (p) (void)hipFree(p);
  delete mg;
  *pmg = nullptr;
  return PTQ2_OK;
}

int ptq2_mg_create(ptq2_mg **out, ptq2_ctx *fine, const ptq2_mg_opts *opts) {
  if (!out || !fine) return PTQ2_ERR_INVALID_ARG;
  *out = nullptr;
  if (!fine->mesh_ready || !fine->eta_ready) return PTQ2_ERR_NOT_SETUP;
  if (fine->box.i0 || fine->box.j0 || fine->box.k0 || fine->box.NX != 2 * fine->box.mx + 1 ||
      fine->box.NY != 2 * fine->box.my + 1 || fine->box.NZ != 2 * fine->box.mz + 1)
    return PTQ2_ERR_UNSUPPORTED;
  ptq2_mg_opts o;
  if (opts) o = *opts;
  else ptq2_mg_default_opts(&o);
  if (o.nlevels < 2 || o.smooth_its < 1 || o.esteig_its < 1 || !(o.emax_frac > 0) || !(o.emin_frac > 0) ||
      o.autotune < 0 || o.autotune > 4 || o.galerkin_levels < 0 || o.galerkin_levels > o.nlevels - 2 ||
      o.coarse_eta_mean < PTQ2_ETA_MEAN_ARITHMETIC || o.coarse_eta_mean > PTQ2_ETA_MEAN_Q1PROJ)
    return PTQ2_ERR_INVALID_ARG;
  if (o.rigid_deflation &&
      (!(o.rigid_min_contrast >= 1.0) || o.rigid_max_inclusions < 1 || o.rigid_max_inclusions > RG_MAXINC))
    return PTQ2_ERR_INVALID_ARG;
  for (int l = 1; l < o.nlevels; l++) {
    int f = 1 << l;
    if (fine->box.mx % f || fine->box.my % f || fine->box.mz % f) return PTQ2_ERR_INVALID_ARG;
  }
  ptq2_mg *mg = new (std::nothrow) ptq2_mg();
  if (!mg) return PTQ2_ERR_ALLOC;
  std::unique_ptr<ptq2_mg, int (*)(ptq2_mg *)> guard(mg, [](ptq2_mg *p) { return ptq2_mg_destroy(&p); });
  mg->opt = o;
  mg->fine_gen = fine->gen;
  mg->nl = o.nlevels;
  mg->f32 = o.fp32_smoother != 0;
  mg->basis32 = o.fp32_basis != 0;
  // set-up and graph capture run on the fine context's own (capturable, non-null) stream
  mg->fine_stream0 = (void *)fine->stream;
  mg->stream = fine->own_stream;
  ptq2_set_stream(fine, (void *)mg->stream);
  MGCHK(hipSetDevice(fine->device));
  MGCHK(hipStreamSynchronize((hipStream_t)mg->fine_stream0));

  // ---- level meshes
  mg->ctx.push_back(fine);
  mg->owned.push_back(false);
  mg->m = {fine->box.mx, fine->box.my, fine->box.mz};
  mg->n.push_back(fine->nvel);
  std::vector<double> cf((size_t)fine->nvel);
  std::vector<uint8_t> mf((size_t)fine->nvel, 0);
  MGCHK(hipMemcpy(cf.data(), fine->coords, sizeof(double) * fine->nvel, hipMemcpyDeviceToHost));
  if (fine->mask) MGCHK(hipMemcpy(mf.data(), fine->mask, fine->nvel, hipMemcpyDeviceToHost));
  std::vector<double> q1nodal;  // lattice field of the reference's coefficient hierarchy
  if (o.coarse_eta_mean == PTQ2_ETA_MEAN_Q1PROJ) {
    std::vector<double> ef((size_t)27 * fine->nel);
    double *d = nullptr;
    MGRC(mg_alloc(mg, (void **)&d, sizeof(double) * ef.size()));
    hipLaunchKernelGGL(k_eta_from_bm, dim3(nblk(27 * fine->nel)), dim3(TPB), 0, mg->stream, fine->box,
                       (const double *)fine->eta_bm, d);
    hipError_t e1 = hipMemcpyAsync(ef.data(), d, sizeof(double) * ef.size(), hipMemcpyDeviceToHost, mg->stream);
    hipError_t e2 = hipStreamSynchronize(mg->stream);
    (void)hipFree(d);
    MGCHK(e1);
    MGCHK(e2);
    q1_nodal_projection(fine->box.mx, fine->box.my, fine->box.mz, ef.data(), q1nodal);
  }
  for (int l = 1; l < mg->nl; l++) {
    const int *pm = &mg->m[3 * (l - 1)];
    int mx = pm[0] / 2, my = pm[1] / 2, mz = pm[2] / 2;
    int NXf = 2 * pm[0] + 1, NYf = 2 * pm[1] + 1;
    int NX = 2 * mx + 1, NY = 2 * my + 1, NZ = 2 * mz + 1;
    std::vector<double> cc((size_t)3 * NX * NY * NZ);
    std::vector<uint8_t> mc((size_t)3 * NX * NY * NZ);
    for (int k = 0; k < NZ; k++)
      for (int j = 0; j < NY; j++)
        for (int i = 0; i < NX; i++) {
          int64_t a = (int64_t)i + (int64_t)NX * (j + (int64_t)NY * k);
          int64_t b = (int64_t)(2 * i) + (int64_t)NXf * (2 * j + (int64_t)NYf * (2 * k));
          for (int d = 0; d < 3; d++) {
            cc[3 * a + d] = cf[3 * b + d];
            mc[3 * a + d] = mf[3 * b + d];
          }
        }
    ptq2_ctx *c = nullptr;
    MGRC(ptq2_create(&c, fine->device));
    mg->ctx.push_back(c);
    mg->owned.push_back(true);
    ptq2_set_stream(c, mg->stream);
    MGRC(ptq2_setup_mesh(c, mx, my, mz, cc.data(), nullptr, fine->mask ? mc.data() : nullptr));
    if (o.coarse_eta_mean == PTQ2_ETA_MEAN_Q1PROJ) {
      // the reference's hierarchy: the lattice field of the level above, restricted, then
      // interpolated to this level's quadrature points
      std::vector<double> nc, ec;
      q1_restrict_nodal(mx, my, mz, q1nodal, nc);
      q1_interpolate_qp(mx, my, mz, nc, ec);
      q1nodal.swap(nc);
      MGRC(ptq2_set_viscosity(c, ec.data(), 0, 0));
    } else {
      // viscosity from the finer level
      double *etac = nullptr;
      MGRC(mg_alloc(mg, (void **)&etac, sizeof(double) * 27 * (size_t)c->nel));
      ptq2_ctx *p = mg->ctx[l - 1];
      hipLaunchKernelGGL(k_eta_coarsen, dim3(nblk(27 * c->nel)), dim3(TPB), 0, mg->stream, p->box, mx, my, mz,
                         (const double *)p->eta_bm, o.coarse_eta_mean, etac);
      int rc = ptq2_set_viscosity(c, etac, 0, PTQ2_DEVICE);
      (void)hipFree(etac);
      MGRC(rc);
    }
    mg->m.insert(mg->m.end(), {mx, my, mz});
    mg->n.push_back(c->nvel);
    cf.swap(cc);
    mf.swap(mc);
  }
  if (mg->n.back() > o.coarse_max_dofs) return mgfail(mg, PTQ2_ERR_SIZE, "coarse level too large for the dense solve");
  mg->tune.assign(mg->nl, LevelTune{});
  for (int l = 0; l < mg->nl; l++) {
    mg->lbox.push_back(mg->ctx[l]->box);
    mg->tune[l].fast = mg->ctx[l]->box.fast32;
  }
  // ---- Galerkin levels: nl-1-galerkin_levels .. nl-2, each the product through the level above
  mg->gal.assign(mg->nl, 0);
  mg->gst64.assign(mg->nl, nullptr);
  mg->gst32.assign(mg->nl, nullptr);
  mg->gmf_in.assign(mg->nl, nullptr);
  mg->gmf_out.assign(mg->nl, nullptr);
  for (int l = 0; l + 1 < mg->nl; l++) {
    int rc = PTQ2_OK;
    (void)ptq2_internal_facebuf(mg->ctx[l], &rc);
    MGRC(rc);
  }
  for (int l = mg->nl - 1 - o.galerkin_levels; l + 1 < mg->nl; l++) MGRC(build_galerkin_level(mg, l));

  // ---- small reduction buffers (also used by the solve)
  MGRC(mg_alloc(mg, (void **)&mg->part, sizeof(double) * (2 * MAXRESTART + 4) * MDB));
  MGRC(mg_alloc(mg, (void **)&mg->psum, sizeof(double) * (2 * MAXRESTART + 4)));
  MGRC(mg_alloc(mg, (void **)&mg->hdev, sizeof(double) * (2 * MAXRESTART + 2)));
  MGRC(mg_alloc(mg, (void **)&mg->Vptr, sizeof(double *) * (MAXRESTART + 1)));
  MGRC(mg_alloc(mg, (void **)&mg->Zptr, sizeof(double *) * (MAXRESTART + 1)));

  // ---- level vectors, Jacobi diagonals, eigen-estimates
  const size_t es = mg->f32 ? sizeof(float) : sizeof(double);
  mg->vb.assign(mg->nl, nullptr); mg->vx.assign(mg->nl, nullptr); mg->vy.assign(mg->nl, nullptr);
  mg->vp1.assign(mg->nl, nullptr); mg->vp2.assign(mg->nl, nullptr); mg->vdinv.assign(mg->nl, nullptr);
  mg->emax.assign(mg->nl, 0.0); mg->emin.assign(mg->nl, 0.0);
  for (int l = 0; l < mg->nl; l++) {
    const int64_t n = mg->n[l];
    MGRC(mg_alloc(mg, &mg->vb[l], es * n));
    MGRC(mg_alloc(mg, &mg->vx[l], es * n));
    if (l == mg->nl - 1) break;
    MGRC(mg_alloc(mg, &mg->vy[l], es * n));
    MGRC(mg_alloc(mg, &mg->vp1[l], es * n));
    MGRC(mg_alloc(mg, &mg->vp2[l], es * n));
    MGRC(mg_alloc(mg, &mg->vdinv[l], es * n));
    ptq2_ctx *c = mg->ctx[l];
    if (mg->f32) MGRC(ptq2_internal_build_fac32(c));
    double *d = nullptr, *v = nullptr, *w = nullptr;
    MGRC(mg_alloc(mg, (void **)&d, sizeof(double) * n));
    MGRC(mg_alloc(mg, (void **)&v, sizeof(double) * n));
    MGRC(mg_alloc(mg, (void **)&w, sizeof(double) * n));
    int rc = PTQ2_OK;
    if (mg->gal[l]) {
      hipLaunchKernelGGL(k_stencil_diag, dim3(nblk(n)), dim3(TPB), 0, mg->stream, c->nnodes,
                         (const double *)mg->gst64[l], d);
      if (hipGetLastError() != hipSuccess) rc = mgfail(mg, PTQ2_ERR_DEVICE, "stencil diagonal");
    } else {
      rc = ptq2_a11_diag(c, d, PTQ2_DEVICE | PTQ2_MATMULT);
    }
    if (rc == PTQ2_OK) {
      if (mg->f32) hipLaunchKernelGGL(k_recip<float>, dim3(nblk(n)), dim3(TPB), 0, mg->stream, n, (const double *)d, (float *)mg->vdinv[l]);
      else hipLaunchKernelGGL(k_recip<double>, dim3(nblk(n)), dim3(TPB), 0, mg->stream, n, (const double *)d, (double *)mg->vdinv[l]);
      // power iteration on D^-1 A (FP64, MatMult semantics) with the Rayleigh quotient
      // v^T A v / v^T D v; the start vector is a SplitMix64 stream
      std::vector<double> hv((size_t)n);
      uint64_t s = 0x9E3779B97F4A7C15ull ^ (uint64_t)l;
      for (int64_t i = 0; i < n; i++) {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        hv[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
      }
      if (hipMemcpy(v, hv.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) rc = PTQ2_ERR_DEVICE;
      double lam = 0.0;
      const unsigned nb = 1024;
      std::vector<double> hp(2 * nb);
      for (int it = 0; it < o.esteig_its && rc == PTQ2_OK; it++) {
        rc = mg->gal[l] ? level_matmult<double>(mg, l, v, w) : ptq2_a11_apply(c, v, w, PTQ2_DEVICE | PTQ2_MATMULT);
        if (rc) break;
        hipLaunchKernelGGL(k_rq_parts, dim3(nb), dim3(TPB), 0, mg->stream, n, (const double *)v, (const double *)w,
                           (const double *)d, mg->part);
        hipLaunchKernelGGL(k_div, dim3(nblk(n)), dim3(TPB), 0, mg->stream, n, (const double *)w, (const double *)d, v);
        if (hipMemcpyAsync(hp.data(), mg->part, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, mg->stream) != hipSuccess ||
            hipStreamSynchronize(mg->stream) != hipSuccess) {
          rc = PTQ2_ERR_DEVICE;
          break;
        }
        double vav = 0.0, vdv = 0.0;
        for (unsigned k = 0; k < nb; k++) {
          vav += hp[2 * k];
          vdv += hp[2 * k + 1];
        }
        lam = vav / vdv;
        double nrm;
        rc = norm2(mg, v, n, &nrm);
        if (rc) break;
        hipLaunchKernelGGL(k_scale_copy, dim3(nblk(n)), dim3(TPB), 0, mg->stream, n, (const double *)v, 1.0 / nrm, v);
      }
      mg->emax[l] = o.emax_frac * lam;
      mg->emin[l] = o.emin_frac * lam;
    }
    (void)hipFree(d); (void)hipFree(v); (void)hipFree(w);
    MGRC(rc);
  }
  // ---- launch shapes of the smoother applies
  for (int l = 0; l + 1 < mg->nl; l++) {
    int rc = PTQ2_OK;
    (void)ptq2_internal_facebuf(mg->ctx[l], &rc);
    MGRC(rc);
  }
  MGRC(mg->f32 ? autotune<float>(mg) : autotune<double>(mg));
  // ---- dense coarse inverse, assembled and inverted on the device
  {
    const int l = mg->nl - 1;
    ptq2_ctx *c = mg->ctx[l];
    const int nc = (int)mg->n[l];
    const int np = (nc + SB - 1) / SB * SB;
    const int nbk = np / SB;
    const int64_t nn = (int64_t)np * np;
    double *e = nullptr, *P = nullptr, *C = nullptr, *Dm = nullptr;
    int *err = nullptr;
    MGRC(mg_alloc(mg, (void **)&mg->ainv, sizeof(double) * (size_t)nn));
    mg->ainv_ld = np;
    MGRC(mg_alloc(mg, (void **)&e, sizeof(double) * nc));
    int rc = PTQ2_OK;
    double *w = nullptr;
    MGRC(mg_alloc(mg, (void **)&w, sizeof(double) * nc));
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(nn)), dim3(TPB), 0, mg->stream, nn, mg->ainv, 0.0);
    hipLaunchKernelGGL(k_pad_identity, dim3((unsigned)((nn + TPB - 1) / TPB)), dim3(TPB), 0, mg->stream, nc, np,
                       mg->ainv);
    const int NXc = 2 * mg->m[3 * l] + 1, NYc = 2 * mg->m[3 * l + 1] + 1, NZc = 2 * mg->m[3 * l + 2] + 1;
    const uint8_t *cmask = c->mask;
    hipLaunchKernelGGL(k_dirichlet_unit_rows, dim3(nblk(nc)), dim3(TPB), 0, mg->stream, (int64_t)nc, cmask, mg->ainv,
                       np);
    // Coloured probing: the coarse operator couples a node only to nodes within PR nodes per axis,
    // so the non-Dirichlet dofs of one component whose node indices agree modulo 2 PR + 1 on every
    // axis have disjoint columns: one probe vector per colour (3 (2 PR + 1)^3 probes instead of
    // one per dof) and k_probe_scatter puts each entry into the row of the probed dof it belongs to.
    //   Galerkin R A_f P: P spreads a coarse node over fine nodes 2I-1..2I+1, A_f reaches two fine
    //   nodes further, R gathers fine nodes 2J-1..2J+1, so |I - J| <= 1 (PR = 1, 81 probes).
    //   Rediscretised Q2 operator: element vertices reach two nodes away (PR = 2, 375 probes).
    const bool gal = mg->opt.galerkin_coarse && mg->nl >= 2;
    const int PR = gal ? 1 : 2, PP = 2 * PR + 1;
    const int lf = l - 1;
    const int64_t nf = gal ? mg->n[lf] : 0;
    double *vf = nullptr, *wf = nullptr;
    if (gal) {
      MGRC(mg_alloc(mg, (void **)&vf, sizeof(double) * nf));
      MGRC(mg_alloc(mg, (void **)&wf, sizeof(double) * nf));
    }
    int probes = 0;
    for (int cz = 0; cz < PP && rc == PTQ2_OK; cz++)
      for (int cy = 0; cy < PP && rc == PTQ2_OK; cy++)
        for (int cx = 0; cx < PP && rc == PTQ2_OK; cx++)
          for (int comp = 0; comp < 3 && rc == PTQ2_OK; comp++) {
            hipLaunchKernelGGL(k_probe_vec, dim3(nblk(nc)), dim3(TPB), 0, mg->stream, NXc, NYc, NZc, PP, cx, cy, cz,
                               comp, cmask, e);
            if (gal) {
              // Pi R A_f P Pi e (Pi zeroes coarse Dirichlet dofs: the probe holds none, the
              // restriction writes zeros there)
              const int *mf = &mg->m[3 * lf];
              const int NXf = 2 * mf[0] + 1, NYf = 2 * mf[1] + 1, NZf = 2 * mf[2] + 1;
              hipLaunchKernelGGL(k_fill<double>, dim3(nblk(nf)), dim3(TPB), 0, mg->stream, nf, vf, 0.0);
              hipLaunchKernelGGL(k_prolong_add<double>, dim3(nblk(nf)), dim3(TPB), 0, mg->stream, NXc, NYc, NXf, NYf,
                                 NZf, (const double *)e, (const uint8_t *)mg->ctx[lf]->mask, vf);
              rc = level_matmult<double>(mg, lf, vf, wf);
              hipLaunchKernelGGL(k_restrict<double>, dim3(nblk(nc)), dim3(TPB), 0, mg->stream, NXc, NYc, NZc, NXf,
                                 NYf, NZf, (const double *)wf, (const double *)nullptr, cmask, w);
            } else {
              rc = ptq2_a11_apply(c, e, w, PTQ2_DEVICE | PTQ2_MATMULT);
            }
            hipLaunchKernelGGL(k_probe_scatter, dim3(nblk(nc)), dim3(TPB), 0, mg->stream, NXc, NYc, NZc, PP, cx, cy,
                               cz, comp, cmask, (const double *)w, mg->ainv, np);
            // bound the work queued ahead of the host
            if (++probes % 32 == 0 && hipStreamSynchronize(mg->stream) != hipSuccess) rc = PTQ2_ERR_DEVICE;
          }
    if (vf) (void)hipFree(vf);
    if (wf) (void)hipFree(wf);
    (void)hipFree(w);
    (void)hipFree(e);
    MGRC(rc);
    hipLaunchKernelGGL(k_symmetrise, dim3((unsigned)((nn + TPB - 1) / TPB)), dim3(TPB), 0, mg->stream, np, mg->ainv);
    MGCHK(hipGetLastError());
    MGRC(mg_alloc(mg, (void **)&P, sizeof(double) * (size_t)np * SB));
    MGRC(mg_alloc(mg, (void **)&C, sizeof(double) * (size_t)np * SB));
    MGRC(mg_alloc(mg, (void **)&Dm, sizeof(double) * SB * SB));
    MGRC(mg_alloc(mg, (void **)&err, sizeof(int)));
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(TPB), 0, mg->stream, (int64_t)1, err, 0);
    for (int k = 0; k < nbk; k++) {
      hipLaunchKernelGGL(k_sweep_diag, dim3(1), dim3(256), 0, mg->stream, np, k, (const double *)mg->ainv, Dm, err);
      hipLaunchKernelGGL(k_sweep_panel, dim3((np + 3) / 4), dim3(256), 0, mg->stream, np, k, (const double *)mg->ainv,
                         (const double *)Dm, P, C);
      hipLaunchKernelGGL(k_sweep_update, dim3(nbk, nbk), dim3(256), 0, mg->stream, np, k, mg->ainv,
                         (const double *)P, (const double *)C, (const double *)Dm);
      MGCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_negate, dim3((unsigned)((nn + TPB - 1) / TPB)), dim3(TPB), 0, mg->stream, nn, mg->ainv);
    int herr = 0;
    MGCHK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, mg->stream));
    MGCHK(hipStreamSynchronize(mg->stream));
    (void)hipFree(P); (void)hipFree(C); (void)hipFree(Dm); (void)hipFree(err);
    if (herr) return mgfail(mg, PTQ2_ERR_MESH, "coarse operator is not positive definite");
  }
  // ---- two-level correction for stiff inclusions
  if (o.rigid_deflation) MGRC(rigid_setup(mg));
  // ---- graph of one V-cycle
  MGRC(mg_alloc(mg, (void **)&mg->io, sizeof(VcIO)));
  if (o.use_graph) {
    MGCHK(hipStreamBeginCapture(mg->stream, hipStreamCaptureModeThreadLocal));
    int rc = vcycle_body(mg);
    hipGraph_t gph = nullptr;
    hipError_t e = hipStreamEndCapture(mg->stream, &gph);
    if (rc) {
      if (gph) (void)hipGraphDestroy(gph);
      return rc;
    }
    if (e != hipSuccess) return mgfail(mg, PTQ2_ERR_DEVICE, hipGetErrorString(e));
    mg->graph = gph;
    MGCHK(hipGraphInstantiate(&mg->exec, mg->graph, nullptr, nullptr, 0));
  }
  MGCHK(hipStreamSynchronize(mg->stream));
  guard.release();
  *out = mg;
  return PTQ2_OK;
}

int ptq2_mg_levels(const ptq2_mg *mg, int *nl, int *m) {
  if (!mg || !nl) return PTQ2_ERR_INVALID_ARG;
  *nl = mg->nl;
  if (m) std::memcpy(m, mg->m.data(), sizeof(int) * mg->m.size());
  return PTQ2_OK;
}

int ptq2_mg_rigid_info(const ptq2_mg *mg, int *ninclusions, int *ncolumns, int32_t *node_label) {
  if (!mg) return PTQ2_ERR_INVALID_ARG;
  if (ninclusions) *ninclusions = mg->rg_nc;
  if (ncolumns) *ncolumns = mg->rg_kept;
  if (node_label) {
    const int64_t nn = mg->ctx[0]->nnodes;
    for (int64_t i = 0; i < nn; i++) node_label[i] = mg->rg_label.empty() ? -1 : mg->rg_label[i];
  }
  return PTQ2_OK;
}

int ptq2_mg_eigs(const ptq2_mg *mg, double *emin, double *emax) {
  if (!mg) return PTQ2_ERR_INVALID_ARG;
  for (int l = 0; l + 1 < mg->nl; l++) {
    if (emin) emin[l] = mg->emin[l];
    if (emax) emax[l] = mg->emax[l];
  }
  return PTQ2_OK;
}

int ptq2_mg_tuning(const ptq2_mg *mg, int *kernel, int *fast_tiles, double *ms, double *ms_default) {
  if (!mg) return PTQ2_ERR_INVALID_ARG;
  for (int l = 0; l + 1 < mg->nl; l++) {
    if (kernel)
      kernel[l] = mg->gal[l] ? (mg->tune[l].split == 4 ? PTQ2_MG_KERNEL_GALERKIN_MF : PTQ2_MG_KERNEL_STENCIL)
                             : mg->tune[l].split;
    if (fast_tiles) fast_tiles[l] = mg->tune[l].fast;
    if (ms) ms[l] = mg->tune[l].ms;
    if (ms_default) ms_default[l] = mg->tune[l].ms_default;
  }
  return PTQ2_OK;
}

int ptq2_mg_set_stream(ptq2_mg *mg, void *s) {
  if (!mg) return PTQ2_ERR_INVALID_ARG;
  mg->stream = (hipStream_t)s;
  for (ptq2_ctx *c : mg->ctx) ptq2_set_stream(c, s);
  return PTQ2_OK;
}

namespace {
// the hierarchy holds device pointers of the fine context's mesh (in its graph) and coarse
// data derived from its viscosity: both must be the ones it was built from
int check_fine(ptq2_mg *mg) {
  if (mg->ctx[0]->gen != mg->fine_gen)
    return mgfail(mg, PTQ2_ERR_NOT_SETUP,
                  "the fine context's mesh or viscosity changed after ptq2_mg_create: rebuild the hierarchy");
  return PTQ2_OK;
}
}  // namespace

int ptq2_mg_vcycle(ptq2_mg *mg, const double *r, double *z) {
  if (!mg || !r || !z) return PTQ2_ERR_INVALID_ARG;
  MGRC(check_fine(mg));
  return vcycle_scaled(mg, r, 1.0, z);
}

int ptq2_mg_level_times(ptq2_mg *mg, const double *r, double *z, int reps, double *ms) {
  if (!mg || !r || !z || !ms || reps < 1) return PTQ2_ERR_INVALID_ARG;
  MGRC(check_fine(mg));
  if (mg->pev.empty()) {
    mg->pev.assign(4 * (size_t)mg->nl, nullptr);
    for (hipEvent_t &e : mg->pev) MGCHK(hipEventCreate(&e));
  }
  std::vector<double> acc((size_t)mg->nl, 0.0);
  hipLaunchKernelGGL(k_set_io, dim3(1), dim3(64), 0, mg->stream, mg->io, r, z, 1.0, 1.0);
  MGCHK(hipGetLastError());
  for (int it = 0; it < reps; it++) {
    mg->prof = true;
    const int rc = vcycle_body(mg);
    mg->prof = false;
    MGRC(rc);
    MGCHK(hipStreamSynchronize(mg->stream));
    for (int l = 0; l < mg->nl; l++) {
      float a = 0.0f, b = 0.0f;
      hipEvent_t *e = &mg->pev[4 * (size_t)l];
      if (l == mg->nl - 1) {
        MGCHK(hipEventElapsedTime(&a, e[0], e[3]));
      } else {
        MGCHK(hipEventElapsedTime(&a, e[0], e[1]));
        MGCHK(hipEventElapsedTime(&b, e[2], e[3]));
      }
      acc[l] += (double)a + (double)b;
    }
  }
  for (int l = 0; l < mg->nl; l++) ms[l] = acc[l] / reps;
  return PTQ2_OK;
}

int ptq2_mg_solve(ptq2_mg *mg, const double *b, double *x, double rtol, int max_it, int restart, int *iters,
                  double *final_rnorm, double *rnorm0) {
  if (!mg || !b || !x || restart < 1 || restart > MAXRESTART || max_it < 1) return PTQ2_ERR_INVALID_ARG;
  MGRC(check_fine(mg));
  const int64_t n = mg->n[0];
  ptq2_ctx *fc = mg->ctx[0];
  // workspace
  const bool b32 = mg->basis32;
  while ((int)mg->Vk.size() < (b32 ? 2 : restart + 1)) {
    double *p = nullptr;
    MGRC(mg_alloc(mg, (void **)&p, sizeof(double) * n));
    mg->Vk.push_back(p);
  }
  while (b32 && (int)mg->V32.size() < restart) {
    float *p = nullptr;
    MGRC(mg_alloc(mg, (void **)&p, sizeof(float) * n));
    mg->V32.push_back(p);
  }
  while ((int)mg->Zk.size() < restart) {
    double *p = nullptr;
    MGRC(mg_alloc(mg, (void **)&p, sizeof(double) * n));
    mg->Zk.push_back(p);
  }
  if (!mg->w) MGRC(mg_alloc(mg, (void **)&mg->w, sizeof(double) * n));
  // candidate j (FP64) and finished basis vector j (basis precision); the pointer table lists
  // the finished vectors
  auto cand = [&](int j) { return b32 ? mg->Vk[j & 1] : mg->Vk[j]; };
  auto fin = [&](int j) -> void * { return b32 ? (void *)mg->V32[j] : (void *)mg->Vk[j]; };
  std::vector<const void *> vp;
  if (b32) vp.assign(mg->V32.begin(), mg->V32.end());
  else vp.assign(mg->Vk.begin(), mg->Vk.end());
  std::vector<const double *> zp(mg->Zk.begin(), mg->Zk.end());
  MGCHK(hipMemcpy((void *)mg->Vptr, vp.data(), sizeof(void *) * vp.size(), hipMemcpyHostToDevice));
  MGCHK(hipMemcpy((void *)mg->Zptr, zp.data(), sizeof(double *) * zp.size(), hipMemcpyHostToDevice));
  ptq2_set_stream(fc, mg->stream);
  mg->hist.clear();
  const unsigned gb = nblk(n);
  hipLaunchKernelGGL(k_fill<double>, dim3(gb), dim3(TPB), 0, mg->stream, n, x, 0.0);
  double bnorm = 0.0;
  MGRC(norm2(mg, b, n, &bnorm));
  if (rnorm0) *rnorm0 = bnorm;
  // device vectors are not scanned element by element: a NaN or Inf anywhere in b shows in its norm
  if (!std::isfinite(bnorm)) return mgfail(mg, PTQ2_ERR_NONFINITE, "non-finite value in the right-hand side");
  int it = 0;
  double rn = bnorm;
  mg->hist.push_back(rn);
  if (bnorm == 0.0) {
    if (iters) *iters = 0;
    if (final_rnorm) *final_rnorm = 0.0;
    return PTQ2_OK;
  }
  const double target = rtol * bnorm;
  // Arnoldi with delayed reorthogonalisation.  Slot j of V holds the candidate u_j unnormalised
  // (u_j = su * slot j) until step j finishes it: one pass computes its reorthogonalisation
  // coefficients a = V_{j-1}^T u_j together with the projection b = V_{j-1}^T w of w = A z_j,
  // and a second pass writes v_j = (u_j - V a) / alpha over the slot and the next candidate
  // w - V b - h_jj v_j into slot j+1.  Column j-1 of H becomes final in step j (the
  // reorthogonalisation moves H(j, j-1) u_j into V a + alpha v_j), so the residual estimate
  // trails the basis by one step.  z_j = M^-1 u_j: FGMRES only needs A Z = V H for the stored
  // Z, so the preconditioner may see the candidate before its reorthogonalisation.
  std::vector<double> H((size_t)(restart + 1) * restart), cs(restart), sn(restart), g(restart + 1), y(restart),
      dots(2 * (size_t)restart + 4), ab(2 * (size_t)restart + 2), av(restart + 1);
  auto Hc = [&](int i, int j) -> double & { return H[(size_t)i * restart + j]; };
  // rotate the finished column j, extend the rotations, return the residual estimate
  auto givens = [&](int j) {
    for (int i = 0; i < j; i++) {
      const double t = cs[i] * Hc(i, j) + sn[i] * Hc(i + 1, j);
      Hc(i + 1, j) = -sn[i] * Hc(i, j) + cs[i] * Hc(i + 1, j);
      Hc(i, j) = t;
    }
    const double den = std::hypot(Hc(j, j), Hc(j + 1, j));
    cs[j] = den > 0 ? Hc(j, j) / den : 1.0;
    sn[j] = den > 0 ? Hc(j + 1, j) / den : 0.0;
    Hc(j, j) = den;
    Hc(j + 1, j) = 0.0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    return std::fabs(g[j + 1]);
  };
  const double **Vp = mg->Vptr;
  // Stop at the rounding floor: a cycle whose estimate reached the target while its true
  // residual fell by less than 10%, and stays at least 1.5x the target, means the FP64
  // residual cannot resolve the target (the estimate and the true residual part once the
  // update falls below the rounding of A x).  A claim missed by a few percent near the target,
  // or one after which the true residual still falls, lets the solve go on.
  bool prev_claimed = false;
  double prev_beta = 0.0;
  mg->stagnated = false;
  while (it < max_it) {
    // r = b - A x, the unnormalised first candidate
    MGRC(ptq2_a11_apply(fc, x, mg->w, PTQ2_DEVICE | PTQ2_MATMULT));
    hipLaunchKernelGGL(k_sub, dim3(gb), dim3(TPB), 0, mg->stream, n, b, (const double *)mg->w, cand(0));
    double beta;
    MGRC(norm2(mg, cand(0), n, &beta));
    if (beta <= target) {
      rn = beta;
      break;
    }
    if (prev_claimed && beta > 1.5 * target && beta > 0.9 * prev_beta) {
      mg->stagnated = true;
      break;
    }
    prev_beta = beta;
    std::fill(g.begin(), g.end(), 0.0);
    std::fill(H.begin(), H.end(), 0.0);
    g[0] = beta;
    double su = 1.0 / beta;  // u_j = su * slot j
    double bprev = 0.0;      // H(j, j-1) before the reorthogonalisation of u_j
    int k = 0;               // finished columns of this cycle
    for (int j = 0; j < restart && it < max_it; j++) {
      MGRC(vcycle_scaled(mg, cand(j), su, mg->Zk[j]));
      MGRC(ptq2_a11_apply(fc, mg->Zk[j], mg->w, PTQ2_DEVICE | PTQ2_MATMULT));
      it++;
      MGRC(mdot2(mg, j, cand(j), mg->w, n, dots.data()));
      double aa = 0.0, abd = 0.0;
      for (int v = 0; v < j; v++) {
        av[v] = su * dots[2 * v];
        aa += av[v] * av[v];
        abd += av[v] * dots[2 * v + 1];
      }
      const double c = su * dots[2 * j], d = su * su * dots[2 * j + 1];
      const double alpha2 = d - aa;
      double alpha, hjj;
      if (alpha2 > 0.5 * d) {
        alpha = std::sqrt(alpha2);
        hjj = (c - abd) / alpha;
      } else {
        // u_j lost more than half its norm to V: finish it explicitly and project w by
        // plain CGS2 against the finished basis
        hipLaunchKernelGGL(k_scale_copy, dim3(gb), dim3(TPB), 0, mg->stream, n, (const double *)cand(j), su, cand(j));
        if (j > 0) MGCHK(hipMemcpyAsync(mg->hdev, av.data(), sizeof(double) * j, hipMemcpyHostToDevice, mg->stream));
        MGRC(maxpy_norm(mg, j, cand(j), n, &alpha));
        if (!(alpha > 0.0)) return mgfail(mg, PTQ2_ERR_DEVICE, "FGMRES: Krylov candidate vanished");
        hipLaunchKernelGGL(k_scale_copy, dim3(gb), dim3(TPB), 0, mg->stream, n, (const double *)cand(j), 1.0 / alpha,
                           cand(j));
        if (b32)
          hipLaunchKernelGGL(k_to_f32, dim3(gb), dim3(TPB), 0, mg->stream, n, (const double *)cand(j), mg->V32[j]);
        hjj = std::nan("");
      }
      // column j-1 is now final
      if (j > 0) {
        for (int i = 0; i < j; i++) Hc(i, j - 1) += bprev * av[i];
        Hc(j, j - 1) = bprev * alpha;
        rn = givens(j - 1);
        mg->hist.push_back(rn);
        k = j;
        if (rn <= target) break;  // z_j and w are not needed
      }
      double bn;
      if (!std::isnan(hjj)) {
        const double gam = hjj / alpha;
        for (int v = 0; v < j; v++) {
          Hc(v, j) = dots[2 * v + 1];
          ab[v] = av[v];
          ab[j + v] = dots[2 * v + 1] - gam * av[v];
        }
        Hc(j, j) = hjj;
        if (j > 0) MGCHK(hipMemcpyAsync(mg->hdev, ab.data(), sizeof(double) * 2 * j, hipMemcpyHostToDevice, mg->stream));
        MGRC(dcgs_update(mg, j, su, 1.0 / alpha, gam, cand(j), fin(j), cand(j + 1), n, &bn));
      } else {
        // CGS with the DGKS second pass against v_0..v_j
        std::vector<double> h(j + 2), h2(j + 2);
        MGRC(mdot(mg, Vp, j + 1, mg->w, n, true, h.data()));
        const double wn = std::sqrt(std::max(h[j + 1], 0.0));
        MGCHK(hipMemcpyAsync(mg->hdev, h.data(), sizeof(double) * (j + 1), hipMemcpyHostToDevice, mg->stream));
        MGRC(maxpy_norm(mg, j + 1, mg->w, n, &bn));
        if (bn < 0.7071067811865476 * wn) {
          MGRC(mdot(mg, Vp, j + 1, mg->w, n, false, h2.data()));
          MGCHK(hipMemcpyAsync(mg->hdev, h2.data(), sizeof(double) * (j + 1), hipMemcpyHostToDevice, mg->stream));
          MGRC(maxpy_norm(mg, j + 1, mg->w, n, &bn));
          for (int i = 0; i <= j; i++) h[i] += h2[i];
        }
        for (int i = 0; i <= j; i++) Hc(i, j) = h[i];
        MGCHK(hipMemcpyAsync(cand(j + 1), mg->w, sizeof(double) * n, hipMemcpyDeviceToDevice, mg->stream));
      }
      Hc(j + 1, j) = bn;
      bprev = bn;
      if (bn == 0.0) {  // the Krylov space is invariant: column j is final
        rn = givens(j);
        mg->hist.push_back(rn);
        k = j + 1;
        break;
      }
      su = 1.0 / bn;
      if (j == restart - 1 || it >= max_it) {
        // end of the cycle: finish column j with the reorthogonalisation of u_{j+1}
        MGRC(mdot(mg, Vp, j + 1, cand(j + 1), n, true, dots.data()));
        double a2 = 0.0;
        for (int v = 0; v <= j; v++) {
          const double a = su * dots[v];
          Hc(v, j) += bprev * a;
          a2 += a * a;
        }
        Hc(j + 1, j) = bprev * std::sqrt(std::max(su * su * dots[j + 1] - a2, 0.0));
        rn = givens(j);
        mg->hist.push_back(rn);
        k = j + 1;
      }
    }
    prev_claimed = rn <= target;
    // x += Z y, with H(0:k,0:k) y = g
    for (int i = k - 1; i >= 0; i--) {
      double s = g[i];
      for (int q = i + 1; q < k; q++) s -= Hc(i, q) * y[q];
      y[i] = s / Hc(i, i);
    }
    if (k > 0) {
      MGCHK(hipMemcpyAsync(mg->hdev, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, mg->stream));
      hipLaunchKernelGGL(k_mcombine, dim3(gb), dim3(TPB), 0, mg->stream, n, mg->Zptr, k, (const double *)mg->hdev, x);
    }
    // the next pass computes the true residual and stops there if it meets the target,
    // else restarts from x
  }
  // true residual
  MGRC(ptq2_a11_apply(fc, x, mg->w, PTQ2_DEVICE | PTQ2_MATMULT));
  hipLaunchKernelGGL(k_sub, dim3(gb), dim3(TPB), 0, mg->stream, n, b, (const double *)mg->w, mg->w);
  double tr;
  MGRC(norm2(mg, mg->w, n, &tr));
  if (iters) *iters = it;
  if (final_rnorm) *final_rnorm = tr;
  return PTQ2_OK;
}

const char *ptq2_mg_last_error(const ptq2_mg *mg) { return mg ? mg->last.c_str() : ""; }

int ptq2_mg_stagnated(const ptq2_mg *mg) { return mg && mg->stagnated ? 1 : 0; }

int ptq2_mg_history(const ptq2_mg *mg, double *hist, int cap, int *n) {
  if (!mg || !n) return PTQ2_ERR_INVALID_ARG;
  int k = (int)mg->hist.size();
  *n = k;
  if (hist)
    for (int i = 0; i < std::min(k, cap); i++) hist[i] = mg->hist[i];
  return PTQ2_OK;
}

}  // extern "C"
