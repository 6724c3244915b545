// ptq2_kernels.h -- host-side launchers of the operator kernels (C++ linkage).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptq2 {
struct Box;  // defined in ptq2_device.h, which every includer includes first
template <typename T>
struct ChebOut;

// Several problems (boxes) in one launch of the A11 kernel and one of the face pass.  Problem p
// takes blocks start[p] .. start[p + 1] - 1 (start[p] a multiple of 8, at least its brick
// count); fac is eta_bm for FP64 (geometry from coords) and the FP32 factors for FP32.
constexpr int BATCH_MAX = 16;
template <typename T>
struct BatchArgs {
  int n, add;
  int start[BATCH_MAX + 1];
  Box box[BATCH_MAX];
  const T *u[BATCH_MAX];
  T *y[BATCH_MAX];
  T *face[BATCH_MAX];
  const uint8_t *mask[BATCH_MAX];
  const double *coords[BATCH_MAX];
  const T *fac[BATCH_MAX];
};
template <typename T>
hipError_t launch_a11_batch(const BatchArgs<T> &a, hipStream_t s);

// error bits raised by validation kernels into a device int
constexpr int ERRBIT_MESH = 1;
constexpr int ERRBIT_VISCOSITY = 2;
constexpr int ERRBIT_NONFINITE = 4;

// A11 launchers: with facebuf (FACEVALS values per brick) all bricks run in one launch,
// storing the nodes they own, and k_gather_faces adds the shared face values into their
// owners (gather = false leaves that to the caller's next kernel, which then reads the face
// slots itself); without facebuf, 8 coloured launches store/add directly.
// co (one-launch path only): the owned nodes that no face slot adds to get the Chebyshev update
// of co instead of being stored into y (the face pass updates the rest)
hipError_t launch_a11_fp64(const Box &b, const double *coords, const double *u, const uint8_t *mask,
                           const double *eta_bm, double *y, bool add, hipStream_t s, double *facebuf = nullptr,
                           bool gather = true, const ChebOut<double> *co = nullptr);
// the full block [A11 A12; A21 0] in one element pass (FP64); fp (27 per element, may be null)
// subtracts the continuity source from yp
hipError_t launch_stokes_fp64(const Box &b, const double *coords, const double *u, const double *p,
                              const uint8_t *mask, const double *eta_bm, double *yu, double *yp, const double *fp,
                              bool add, hipStream_t s, double *facebuf = nullptr);
hipError_t launch_a11_fp32(const Box &b, const float *u, const uint8_t *mask, const float *fac, float *y, bool add,
                           hipStream_t s, float *facebuf = nullptr, bool gather = true,
                           const ChebOut<float> *co = nullptr);
// Row-split variants (3 waves per brick, one slab of quadrature rows each) for levels with few
// bricks; one-launch output path only (facebuf required)
hipError_t launch_a11_fp64_split(const Box &b, const double *coords, const double *u, const uint8_t *mask,
                                 const double *eta_bm, double *y, bool add, hipStream_t s, double *facebuf,
                                 bool gather = true);
hipError_t launch_a11_fp32_split(const Box &b, const float *u, const uint8_t *mask, const float *fac, float *y,
                                 bool add, hipStream_t s, float *facebuf, bool gather = true);
// The FP32 smoother operator with the reference gradient as a dense GEMM on the matrix cores
// (ptq2_mfma.hip); one-launch output path only
hipError_t launch_a11_fp32_mfma(const Box &b, const float *u, const uint8_t *mask, const float *fac, float *y,
                                bool add, hipStream_t s, float *facebuf, bool gather = true);
// the face pass alone (adds the face slots of facebuf into their owners)
hipError_t launch_gather_faces(const Box &b, const float *face, float *y, hipStream_t s);
hipError_t launch_gather_faces(const Box &b, const double *face, double *y, hipStream_t s);
hipError_t launch_build_fac32(const Box &b, const double *coords, const double *eta_bm, float *fac, float *facb,
                              int *err, hipStream_t s);
hipError_t launch_a11_fp64g(const Box &b, const double *u, const uint8_t *mask, const double *geo, const double *eta_bm,
                            double *y, bool add, hipStream_t s, double *facebuf = nullptr);
hipError_t launch_build_geo64(const Box &b, const double *coords, double *geo, hipStream_t s);
hipError_t launch_stokes_fp32(const Box &b, const float *u, const float *p, const uint8_t *mask, const float *fac,
                              const float *facb, float *yu, float *yp, bool add, hipStream_t s,
                              float *facebuf = nullptr);
hipError_t launch_eta_to_bm(const Box &b, const double *eta, double *eta_bm, int *err, hipStream_t s);
hipError_t launch_check_mesh(const Box &b, const double *coords, int *err, hipStream_t s);
hipError_t launch_a21(const Box &b, const double *coords, const double *u, const uint8_t *mask, double *yp, bool add,
                      hipStream_t s);
hipError_t launch_a12(const Box &b, const double *coords, const double *p, double *y, bool add, hipStream_t s);
hipError_t launch_diag(const Box &b, const double *coords, const double *eta_bm, double *d, bool add, hipStream_t s);
// y (+)= the body-force load of fu (3 per quadrature point, element-major)
// one launch with owner-writes output into facebuf (nbricks * FACEVALS doubles) plus the face pass
hipError_t launch_body_force(const Box &b, const double *coords, const double *fu, double *y, bool add,
                             double *facebuf, hipStream_t s, double scale = 1.0);
template <typename T>
hipError_t launch_bc_rows(int64_t n, const int64_t *idx, const T *x, T *y, int mode, hipStream_t s,
                          const T *v = nullptr);
template <typename T>
hipError_t launch_scan_finite(int64_t n, const T *x, int *err, hipStream_t s);
}  // namespace ptq2
