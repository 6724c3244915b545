// ptq2_kernels.h -- host-side launchers of the operator kernels (C++ linkage).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptq2 {
struct Box;
template <typename T>
struct ChebOut;

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
hipError_t launch_body_force(const Box &b, const double *coords, const double *fu, double *y, bool add,
                             hipStream_t s, double scale = 1.0);
template <typename T>
hipError_t launch_bc_rows(int64_t n, const int64_t *idx, const T *x, T *y, int mode, hipStream_t s,
                          const T *v = nullptr);
template <typename T>
hipError_t launch_scan_finite(int64_t n, const T *x, int *err, hipStream_t s);
}  // namespace ptq2
