// ptq2_points.hip -- material points -> quadrature points on the device.
//
// The Q1 projection of SwarmUpdateGaussPropertiesLocalL2Projection_Q1_MPntPStokes
// (src/material_point_utils.c:905-955, body :663-830): every point adds w * value and w into
// the 27 nodes of its element (w = the Q1 basis at the point for the 8 vertex nodes, 1 for the
// other 19), nodal value = sum / weight sum, and each quadrature point takes the Q1
// interpolation of its element's vertex values clipped to the element's nodal range.
//
// The reference scatters point by point into the nodes (an order the device cannot keep
// without atomics).  Here the points are bucketed by element with a stable radix sort (the
// points of one element keep their index order), and every node gathers from its <= 8
// adjacent elements in ascending element order: one thread per node, no atomics, the same
// sums in a fixed order, so the result is bitwise reproducible.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdarg>
#include <cstdio>

#include "ptq2.h"
#include "ptq2_ctx.h"
#include "ptq2_kernels.h"

namespace ptq2 {
namespace {

constexpr int PTB = 256;

inline unsigned grid_of(int64_t n) {
  const int64_t g = (n + PTB - 1) / PTB;
  return (unsigned)(g < 65535 * 64 ? (g > 0 ? g : 1) : 65535 * 64);
}

// element index in range, local coordinates and values finite
__global__ void k_points_check(int64_t np, int64_t ne, const double *__restrict__ xi, const int32_t *__restrict__ wil,
                               const double *__restrict__ eta, const double *__restrict__ rho, int *__restrict__ err) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t e = wil[p];
    int bits = 0;
    if (e < 0 || (int64_t)e >= ne) bits |= 1;
    bool fin = isfinite(xi[3 * p]) && isfinite(xi[3 * p + 1]) && isfinite(xi[3 * p + 2]) && isfinite(eta[p]);
    if (rho) fin = fin && isfinite(rho[p]);
    if (!fin) bits |= 2;
    if (bits) atomicOr(err, bits);
  }
}

__global__ void k_iota(int64_t n, int32_t *__restrict__ v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = (int32_t)i;
}

// first sorted position of every element (keys sorted ascending): off[e] = lower_bound(e)
__global__ void k_bucket_offsets(int64_t np, int64_t ne, const int32_t *__restrict__ keys, int64_t *__restrict__ off) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e <= ne; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = np;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)keys[mid] < e)
        lo = mid + 1;
      else
        hi = mid;
    }
    off[e] = lo;
  }
}

// the elements adjacent to node index I along one axis (m elements) and the node's local
// index in each: even I is a vertex of elements I/2-1 (local 2) and I/2 (local 0), odd I the
// middle node of element (I-1)/2
__device__ __forceinline__ int adjacent(int I, int m, int el[2], int loc[2]) {
  if (I & 1) {
    el[0] = (I - 1) >> 1;
    loc[0] = 1;
    return 1;
  }
  int n = 0;
  if (I / 2 - 1 >= 0) { el[n] = I / 2 - 1; loc[n] = 2; n++; }
  if (I / 2 < m) { el[n] = I / 2; loc[n] = 0; n++; }
  return n;
}

// Q1 basis of the vertex at local node index l (0 or 2) along one axis at coordinate x
__device__ __forceinline__ double q1_1d(int l, double x) { return l == 0 ? 1.0 - x : 1.0 + x; }

// One thread per node: sum w * eta, w * rho and w over the points of the adjacent elements
// (ascending element, then the points of an element in index order), divide.
__global__ void k_points_to_nodes(Box b, const int64_t *__restrict__ off, const int32_t *__restrict__ order,
                                  const double *__restrict__ xi, const double *__restrict__ eta,
                                  const double *__restrict__ rho, double *__restrict__ neta,
                                  double *__restrict__ nrho) {
  const int64_t nn = (int64_t)b.NX * b.NY * b.NZ;
  for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += (int64_t)gridDim.x * blockDim.x) {
    const int I = (int)(n % b.NX), J = (int)((n / b.NX) % b.NY), K = (int)(n / ((int64_t)b.NX * b.NY));
    int ex[2], lx[2], ey[2], ly[2], ez[2], lz[2];
    const int nx = adjacent(I, b.mx, ex, lx), ny = adjacent(J, b.my, ey, ly), nz = adjacent(K, b.mz, ez, lz);
    double a1 = 0.0, a2 = 0.0, bw = 0.0;
    for (int c = 0; c < nz; c++)
      for (int bb = 0; bb < ny; bb++)
        for (int a = 0; a < nx; a++) {
          const int64_t e = (int64_t)ex[a] + (int64_t)b.mx * ((int64_t)ey[bb] + (int64_t)b.my * ez[c]);
          const bool vert = lx[a] != 1 && ly[bb] != 1 && lz[c] != 1;
          for (int64_t k = off[e]; k < off[e + 1]; k++) {
            const int64_t p = order[k];
            double w = 1.0;
            if (vert)
              w = 0.125 * q1_1d(lx[a], xi[3 * p]) * q1_1d(ly[bb], xi[3 * p + 1]) * q1_1d(lz[c], xi[3 * p + 2]);
            a1 += w * eta[p];
            if (rho) a2 += w * rho[p];
            bw += w;
          }
        }
    neta[n] = bw != 0.0 ? a1 / bw : 0.0;
    if (rho) nrho[n] = bw != 0.0 ? a2 / bw : 0.0;
  }
}

// One thread per element: Q1 interpolation of the vertex values at the 27 Gauss points
// (the constants of QuadratureCreateGauss_3pnt_3D, src/quadrature.c:136-167), clipped to the
// range of the element's 27 nodal values.
__device__ __forceinline__ void interp_element(const Box &b, int64_t e, const double *__restrict__ nv,
                                               double *__restrict__ out) {
  const int ex = (int)(e % b.mx), ey = (int)((e / b.mx) % b.my), ez = (int)(e / ((int64_t)b.mx * b.my));
  const int64_t n0 = 2 * (int64_t)ex + (int64_t)b.NX * (2 * (int64_t)ey + (int64_t)b.NY * 2 * (int64_t)ez);
  double lo = 1.0e32, hi = -1.0e32, v[8];
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        const double a = nv[n0 + i + (int64_t)b.NX * (j + (int64_t)b.NY * k)];
        lo = a < lo ? a : lo;
        hi = a > hi ? a : hi;
        if (i != 1 && j != 1 && k != 1) v[(i >> 1) + 2 * (j >> 1) + 4 * (k >> 1)] = a;
      }
  const double s = 0.774596669241483, g[3] = {-s, 0.0, s};
  for (int q = 0; q < 27; q++) {
    const double x = g[q % 3], y = g[(q / 3) % 3], z = g[q / 9];
    double r = 0.0;
    for (int c = 0; c < 8; c++) {
      const double w = 0.125 * ((c & 1) ? 1.0 + x : 1.0 - x) * ((c & 2) ? 1.0 + y : 1.0 - y) *
                       ((c & 4) ? 1.0 + z : 1.0 - z);
      r += w * v[c];
    }
    r = r < lo ? lo : r;
    r = r > hi ? hi : r;
    out[27 * e + q] = r;
  }
}

__global__ void k_nodes_to_qp(Box b, const double *__restrict__ neta, const double *__restrict__ nrho,
                              double *__restrict__ eta_qp, double *__restrict__ rho_qp) {
  const int64_t ne = (int64_t)b.mx * b.my * b.mz;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    interp_element(b, e, neta, eta_qp);
    if (nrho) interp_element(b, e, nrho, rho_qp);
  }
}

int pfail(ptq2_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  c->last = buf;
  return code;
}

// device scratch of one projection call, freed on every return path
struct Scratch {
  static constexpr int CAP = 16;  // a call takes at most 13 buffers
  void *p[CAP] = {};
  int n = 0;
  ~Scratch() {
    for (int i = 0; i < n; i++)
      if (p[i]) (void)hipFree(p[i]);
  }
  template <typename T>
  hipError_t get(T **out, size_t bytes) {
    void *q = nullptr;
    if (n == CAP) {
      *out = nullptr;
      return hipErrorOutOfMemory;
    }
    hipError_t e = hipMalloc(&q, bytes ? bytes : 1);
    if (e == hipSuccess) p[n++] = q;
    *out = (T *)q;
    return e;
  }
};

}  // namespace
}  // namespace ptq2

#define PCHK(call)                                                                                    \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess) {                                                                           \
      (void)hipGetLastError();                                                                        \
      return ptq2::pfail(c, e_ == hipErrorOutOfMemory ? PTQ2_ERR_ALLOC : PTQ2_ERR_DEVICE, "%s: %s", #call, \
                         hipGetErrorString(e_));                                                      \
    }                                                                                                 \
  } while (0)

extern "C" int ptq2_project_points_q1(ptq2_ctx *c, int64_t np, const double *xi, const int32_t *wil,
                                      const double *eta_p, const double *rho_p, double *eta_qp, double *rho_qp,
                                      unsigned flags) {
  using namespace ptq2;
  if (!c) return PTQ2_ERR_INVALID_ARG;
  if (!c->mesh_ready) return pfail(c, PTQ2_ERR_NOT_SETUP, "MF operator not setup: no mesh");
  if (flags & ~PTQ2_DEVICE) return pfail(c, PTQ2_ERR_INVALID_ARG, "bad flags");
  if (np < 0 || (np > 0 && (!xi || !wil || !eta_p)) || !eta_qp || (!rho_p != !rho_qp))
    return pfail(c, PTQ2_ERR_INVALID_ARG, "null point arrays or outputs (rho_p and rho_qp go together)");
  if (np >= ((int64_t)1 << 31)) return pfail(c, PTQ2_ERR_SIZE, "more than 2^31 - 1 points");
  const Box &b = c->box;
  if (b.i0 || b.j0 || b.k0 || b.NX != 2 * b.mx + 1 || b.NY != 2 * b.my + 1 || b.NZ != 2 * b.mz + 1)
    return pfail(c, PTQ2_ERR_UNSUPPORTED, "point projection needs a single-rank box (no ghost layer)");
  PCHK(hipSetDevice(c->device));
  const bool dev = flags & PTQ2_DEVICE;
  const int64_t ne = c->nel, nn = c->nnodes;
  hipStream_t s = c->stream;
  Scratch S;
  const double *dxi = xi, *deta = eta_p, *drho = rho_p;
  const int32_t *dwil = wil;
  double *deq = eta_qp, *drq = rho_qp;
  if (!dev) {
    double *t;
    int32_t *w;
    PCHK(S.get(&t, sizeof(double) * 3 * (size_t)np));
    PCHK(hipMemcpyAsync(t, xi, sizeof(double) * 3 * (size_t)np, hipMemcpyHostToDevice, s));
    dxi = t;
    PCHK(S.get(&w, sizeof(int32_t) * (size_t)np));
    PCHK(hipMemcpyAsync(w, wil, sizeof(int32_t) * (size_t)np, hipMemcpyHostToDevice, s));
    dwil = w;
    PCHK(S.get(&t, sizeof(double) * (size_t)np));
    PCHK(hipMemcpyAsync(t, eta_p, sizeof(double) * (size_t)np, hipMemcpyHostToDevice, s));
    deta = t;
    if (rho_p) {
      PCHK(S.get(&t, sizeof(double) * (size_t)np));
      PCHK(hipMemcpyAsync(t, rho_p, sizeof(double) * (size_t)np, hipMemcpyHostToDevice, s));
      drho = t;
    }
    PCHK(S.get(&deq, sizeof(double) * 27 * (size_t)ne));
    if (rho_qp) PCHK(S.get(&drq, sizeof(double) * 27 * (size_t)ne));
  }
  // validate before anything reads through wil
  PCHK(hipMemsetAsync(c->err, 0, sizeof(int), s));
  if (np > 0) {
    hipLaunchKernelGGL(k_points_check, dim3(grid_of(np)), dim3(PTB), 0, s, np, ne, dxi, dwil, deta, drho, c->err);
    PCHK(hipGetLastError());
  }
  int bits = 0;
  PCHK(hipMemcpyAsync(&bits, c->err, sizeof(int), hipMemcpyDeviceToHost, s));
  PCHK(hipStreamSynchronize(s));
  if (bits & 1) return pfail(c, PTQ2_ERR_INVALID_ARG, "point element index outside [0, %lld)", (long long)ne);
  if (bits & 2) return pfail(c, PTQ2_ERR_NONFINITE, "non-finite point coordinate or value");
  // stable bucketing by element
  int32_t *keys, *order, *iota;
  int64_t *off;
  PCHK(S.get(&keys, sizeof(int32_t) * (size_t)np));
  PCHK(S.get(&order, sizeof(int32_t) * (size_t)np));
  PCHK(S.get(&iota, sizeof(int32_t) * (size_t)np));
  PCHK(S.get(&off, sizeof(int64_t) * (size_t)(ne + 1)));
  if (np > 0) {
    hipLaunchKernelGGL(k_iota, dim3(grid_of(np)), dim3(PTB), 0, s, np, iota);
    PCHK(hipGetLastError());
    int end_bit = 1;
    while (end_bit < 31 && ((int64_t)1 << end_bit) < ne) end_bit++;
    size_t tmp_bytes = 0;
    // the keys are validated non-negative: sort them as unsigned on the bits that can be set
    const uint32_t *ukin = reinterpret_cast<const uint32_t *>(dwil);
    uint32_t *ukout = reinterpret_cast<uint32_t *>(keys);
    PCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, ukin, ukout, iota, order, (int)np, 0, end_bit, s));
    void *tmp;
    PCHK(S.get(&tmp, tmp_bytes));
    PCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, ukin, ukout, iota, order, (int)np, 0, end_bit, s));
  }
  hipLaunchKernelGGL(k_bucket_offsets, dim3(grid_of(ne + 1)), dim3(PTB), 0, s, np, ne, (const int32_t *)keys, off);
  PCHK(hipGetLastError());
  double *neta, *nrho = nullptr;
  PCHK(S.get(&neta, sizeof(double) * (size_t)nn));
  if (drho) PCHK(S.get(&nrho, sizeof(double) * (size_t)nn));
  hipLaunchKernelGGL(k_points_to_nodes, dim3(grid_of(nn)), dim3(PTB), 0, s, b, (const int64_t *)off,
                     (const int32_t *)order, dxi, deta, drho, neta, nrho);
  PCHK(hipGetLastError());
  hipLaunchKernelGGL(k_nodes_to_qp, dim3(grid_of(ne)), dim3(PTB), 0, s, b, (const double *)neta,
                     (const double *)nrho, deq, drq);
  PCHK(hipGetLastError());
  if (!dev) {
    PCHK(hipMemcpyAsync(eta_qp, deq, sizeof(double) * 27 * (size_t)ne, hipMemcpyDeviceToHost, s));
    if (rho_qp) PCHK(hipMemcpyAsync(rho_qp, drq, sizeof(double) * 27 * (size_t)ne, hipMemcpyDeviceToHost, s));
  }
  // the scratch is freed on return: wait for the kernels that read it
  PCHK(hipStreamSynchronize(s));
  c->last.clear();
  return PTQ2_OK;
}
