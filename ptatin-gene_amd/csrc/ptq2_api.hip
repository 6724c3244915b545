// ptq2_api.hip -- C ABI of libptq2 (declared in include/ptq2.h).
//
// A context owns one device-resident mesh: FP64 coordinates in their natural
// interlaced layout, the Dirichlet mask and the list of Dirichlet dofs, the
// quadrature-point viscosity in brick-major order and, on demand, the FP32
// geometric factors of the smoother operator.  Host-pointer calls stage the
// vectors through device buffers and synchronise (the behaviour of the
// reference's CUDA backend, src/stokes_operators_cuda.cu:325-455, minus the
// per-call re-uploads of mesh data); device-pointer calls are stream ordered.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ptq2.h"
#include "ptq2_device.h"
#include "ptq2_kernels.h"
#include "ptq2_ctx.h"

using ptq2::Box;

namespace {

// Face slots of ptq2_a11_apply_noalloc: a library-owned device array instead of a per-context
// buffer, so that the call allocates nothing.  Each GPU holds its own copy from the first use of
// the library on it until the process ends.  Host code never names the array in a launch or copy:
// its host-side symbol is not a device address.  ptq2_create resolves the device address once
// per device (hipGetSymbolAddress with that device current) and the entry point uses that.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;  // 32 MiB per GPU
__device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];
constexpr int NOALLOC_MAXDEV = 64;
std::mutex g_scratch_mu;
double *g_scratch[NOALLOC_MAXDEV] = {};  // device address of `scratch` per device, or null

// the device address of `scratch` on the current device `dev` (resolved on first call)
hipError_t scratch_resolve(int dev, double **out) {
  *out = nullptr;
  if (dev < 0 || dev >= NOALLOC_MAXDEV) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = (double *)ptr;
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

int fail(ptq2_ctx *c, int code, const char *fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->last = buf;
  }
  return code;
}

#define HIPCHK(c, call)                                                                            \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) return fail((c), PTQ2_ERR_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

int dev_alloc(ptq2_ctx *c, void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return PTQ2_OK;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(c, PTQ2_ERR_ALLOC, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
  }
  return PTQ2_OK;
}

void dev_free(void *p) {
  if (p) (void)hipFree(p);
}

int ensure(ptq2_ctx *c, void **buf, size_t *have, size_t need) {
  if (*have >= need) return PTQ2_OK;
  dev_free(*buf);
  *buf = nullptr;
  *have = 0;
  int rc = dev_alloc(c, buf, need);
  if (rc == PTQ2_OK) *have = need;
  return rc;
}

int set_device(ptq2_ctx *c) {
  HIPCHK(c, hipSetDevice(c->device));
  return PTQ2_OK;
}

// clear the device error word, run a validating launch sequence, read it back
int read_err(ptq2_ctx *c, int *bits) {
  HIPCHK(c, hipMemcpyAsync(bits, c->err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PTQ2_OK;
}

int clear_err(ptq2_ctx *c) {
  HIPCHK(c, hipMemsetAsync(c->err, 0, sizeof(int), c->stream));
  return PTQ2_OK;
}

}  // namespace

int ptq2_internal_build_fac32(ptq2_ctx *c, bool block) {
  if (c->fac32_valid && (!block || c->facb32_valid)) return PTQ2_OK;
  if (!c->fac32) {
    int rc = dev_alloc(c, (void **)&c->fac32, sizeof(float) * (size_t)c->nbricks * 243 * ptq2::BEL);
    if (rc) return rc;
  }
  if (block && !c->facb32) {
    int rc = dev_alloc(c, (void **)&c->facb32, sizeof(float) * (size_t)c->nbricks * 108 * ptq2::BEL);
    if (rc) return rc;
  }
  int rc = clear_err(c);
  if (rc) return rc;
  HIPCHK(c, ptq2::launch_build_fac32(c->box, c->coords, c->eta_bm, c->fac32, block ? c->facb32 : nullptr, c->err,
                                     c->stream));
  int bits = 0;
  rc = read_err(c, &bits);
  if (rc) return rc;
  if (bits & ptq2::ERRBIT_MESH) return fail(c, PTQ2_ERR_MESH, "non-positive Jacobian determinant");
  if (bits & ptq2::ERRBIT_VISCOSITY) return fail(c, PTQ2_ERR_VISCOSITY, "eta <= 0 or non-finite");
  c->fac32_valid = true;
  if (block) c->facb32_valid = true;
  return PTQ2_OK;
}

void *ptq2_internal_facebuf(ptq2_ctx *c, int *rc) {
  *rc = PTQ2_OK;
  if (!c->facebuf) *rc = dev_alloc(c, &c->facebuf, sizeof(double) * (size_t)c->nbricks * ptq2::FACEVALS);
  return c->facebuf;
}

namespace {

int build_fac32(ptq2_ctx *c) { return ptq2_internal_build_fac32(c); }

// resident FP64 geometry of the stored-geometry A11 kernel (depends on the mesh only)
int build_geo64(ptq2_ctx *c) {
  if (c->geo64_valid) return PTQ2_OK;
  if (!c->geo64) {
    int rc = dev_alloc(c, (void **)&c->geo64, sizeof(double) * (size_t)c->nbricks * 243 * ptq2::BEL);
    if (rc) return rc;
  }
  HIPCHK(c, ptq2::launch_build_geo64(c->box, c->coords, c->geo64, c->stream));
  c->geo64_valid = true;
  return PTQ2_OK;
}

int check_ready(ptq2_ctx *c, bool need_eta) {
  if (!c) return PTQ2_ERR_INVALID_ARG;
  if (!c->mesh_ready) return fail(c, PTQ2_ERR_NOT_SETUP, "MF operator not setup: no mesh");
  if (need_eta && !c->eta_ready) return fail(c, PTQ2_ERR_NOT_SETUP, "MF operator not setup: no viscosity");
  return PTQ2_OK;
}

// stage an input vector (host path) and scan it for NaN/Inf
template <typename T>
int stage_in(ptq2_ctx *c, const void *h, void **buf, size_t *have, int64_t n, bool scan) {
  int rc = ensure(c, buf, have, sizeof(T) * (size_t)n);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(*buf, h, sizeof(T) * (size_t)n, hipMemcpyHostToDevice, c->stream));
  if (scan) HIPCHK(c, ptq2::launch_scan_finite<T>(n, (const T *)*buf, c->err, c->stream));
  return PTQ2_OK;
}

}  // namespace

extern "C" {

const char *ptq2_strerror(int code) {
  switch (code) {
    case PTQ2_OK: return "success";
    case PTQ2_ERR_INVALID_ARG: return "invalid argument";
    case PTQ2_ERR_NOT_SETUP: return "operator not set up";
    case PTQ2_ERR_SIZE: return "size out of range";
    case PTQ2_ERR_ALLOC: return "allocation failed";
    case PTQ2_ERR_DEVICE: return "HIP runtime error";
    case PTQ2_ERR_NONFINITE: return "non-finite value in input";
    case PTQ2_ERR_MESH: return "non-positive Jacobian determinant";
    case PTQ2_ERR_VISCOSITY: return "non-positive viscosity";
    case PTQ2_ERR_UNSUPPORTED: return "unsupported connectivity";
    default: return "unknown error";
  }
}

const char *ptq2_last_error(const ptq2_ctx *ctx) { return ctx ? ctx->last.c_str() : ""; }

int ptq2_create(ptq2_ctx **out, int device) {
  if (!out) return PTQ2_ERR_INVALID_ARG;
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return PTQ2_ERR_DEVICE;
  }
  if (device < 0 || device >= ndev) return PTQ2_ERR_INVALID_ARG;
  ptq2_ctx *c = new (std::nothrow) ptq2_ctx();
  if (!c) return PTQ2_ERR_ALLOC;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return PTQ2_ERR_DEVICE;
  }
  c->stream = c->own_stream;
  double *scr = nullptr;
  if (scratch_resolve(device, &scr) != hipSuccess) {  // the device is current (hipSetDevice above)
    (void)hipGetLastError();
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return PTQ2_ERR_DEVICE;
  }
  if (hipMalloc((void **)&c->err, sizeof(int)) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return PTQ2_ERR_ALLOC;
  }
  *out = c;
  return PTQ2_OK;
}

int ptq2_destroy(ptq2_ctx **pc) {
  if (!pc || !*pc) return PTQ2_OK;
  ptq2_ctx *c = *pc;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dev_free(c->coords); dev_free(c->mask); dev_free(c->dir_idx); dev_free(c->eta_bm);
  dev_free(c->fac32); dev_free(c->facb32); dev_free(c->geo64); dev_free(c->facebuf);
  dev_free(c->err); dev_free(c->su); dev_free(c->sy); dev_free(c->sp); dev_free(c->syp); dev_free(c->ures);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  if (c->ev) (void)hipEventDestroy(c->ev);
  delete c;
  *pc = nullptr;
  return PTQ2_OK;
}

int ptq2_set_stream(ptq2_ctx *c, void *s) {
  if (!c) return PTQ2_ERR_INVALID_ARG;
  c->stream = (hipStream_t)s;
  return PTQ2_OK;
}

void *ptq2_get_stream(const ptq2_ctx *c) { return c ? (void *)c->stream : nullptr; }

int ptq2_sync(ptq2_ctx *c) {
  if (!c) return PTQ2_ERR_INVALID_ARG;
  int rc = set_device(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PTQ2_OK;
}

int ptq2_sizes(const ptq2_ctx *c, int64_t *nv, int64_t *np, int64_t *ne) {
  if (!c) return PTQ2_ERR_INVALID_ARG;
  if (!c->mesh_ready) return PTQ2_ERR_NOT_SETUP;
  if (nv) *nv = c->nvel;
  if (np) *np = c->npres;
  if (ne) *ne = c->nel;
  return PTQ2_OK;
}

int ptq2_setup_mesh_box(ptq2_ctx *c, int mx, int my, int mz, int NX, int NY, int NZ, int i0, int j0, int k0,
                        const double *coords, const uint8_t *mask) {
  if (!c || !coords) return fail(c, PTQ2_ERR_INVALID_ARG, "null context or coordinates");
  if (mx < 1 || my < 1 || mz < 1 || i0 < 0 || j0 < 0 || k0 < 0)
    return fail(c, PTQ2_ERR_INVALID_ARG, "bad element counts or offsets");
  if ((int64_t)i0 + 2 * (int64_t)mx + 1 > NX || (int64_t)j0 + 2 * (int64_t)my + 1 > NY ||
      (int64_t)k0 + 2 * (int64_t)mz + 1 > NZ)
    return fail(c, PTQ2_ERR_INVALID_ARG, "elements do not fit in the node box");
  const int64_t nn = (int64_t)NX * NY * NZ;
  const int64_t ne = (int64_t)mx * my * mz;
  // Supported maximum: 2^32 nodes (103 GB of coordinates) and 2^31 elements; checked before
  // any caller memory is read.
  if (nn > ((int64_t)1 << 32) || ne > ((int64_t)1 << 31)) return fail(c, PTQ2_ERR_SIZE, "mesh too large");
  int rc = set_device(c);
  if (rc) return rc;
  for (int64_t i = 0; i < 3 * nn; i++)
    if (!std::isfinite(coords[i])) return fail(c, PTQ2_ERR_NONFINITE, "non-finite coordinate at dof %lld", (long long)i);

  // release a previous mesh
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->gen++;
  dev_free(c->coords); dev_free(c->mask); dev_free(c->dir_idx); dev_free(c->eta_bm);
  dev_free(c->fac32); dev_free(c->facb32); dev_free(c->geo64); dev_free(c->facebuf);
  c->coords = nullptr; c->mask = nullptr; c->dir_idx = nullptr; c->eta_bm = nullptr;
  c->fac32 = nullptr; c->facb32 = nullptr; c->geo64 = nullptr; c->facebuf = nullptr;
  c->mesh_ready = c->eta_ready = c->fac32_valid = c->facb32_valid = c->geo64_valid = c->has_state = false;
  c->ndir = 0;

  Box b;
  b.mx = mx; b.my = my; b.mz = mz; b.NX = NX; b.NY = NY; b.NZ = NZ; b.i0 = i0; b.j0 = j0; b.k0 = k0;
  b.nbx = (mx + ptq2::BE - 1) / ptq2::BE; b.nby = (my + ptq2::BE - 1) / ptq2::BE; b.nbz = (mz + ptq2::BE - 1) / ptq2::BE;
  {
    // byte offsets of a 9^3-node tile's values from its first node fit 32 bits (fast tile paths)
    const int64_t sy = 3 * (int64_t)NX, sz = sy * NY;
    b.fast32 = 8 * (3 * (int64_t)ptq2::TN + (ptq2::TN - 1) * (sy + sz)) < ((int64_t)1 << 31) ? 1 : 0;
  }
  const int64_t nb = (int64_t)b.nbx * b.nby * b.nbz;
  if (nb > INT32_MAX / 2) return fail(c, PTQ2_ERR_SIZE, "too many bricks");

  if ((rc = dev_alloc(c, (void **)&c->coords, sizeof(double) * 3 * (size_t)nn))) return rc;
  HIPCHK(c, hipMemcpy(c->coords, coords, sizeof(double) * 3 * (size_t)nn, hipMemcpyHostToDevice));
  if (mask) {
    std::vector<int64_t> idx;
    for (int64_t i = 0; i < 3 * nn; i++)
      if (mask[i]) idx.push_back(i);
    if ((rc = dev_alloc(c, (void **)&c->mask, 3 * (size_t)nn))) return rc;
    HIPCHK(c, hipMemcpy(c->mask, mask, 3 * (size_t)nn, hipMemcpyHostToDevice));
    c->ndir = (int64_t)idx.size();
    if (c->ndir) {
      if ((rc = dev_alloc(c, (void **)&c->dir_idx, sizeof(int64_t) * idx.size()))) return rc;
      HIPCHK(c, hipMemcpy(c->dir_idx, idx.data(), sizeof(int64_t) * idx.size(), hipMemcpyHostToDevice));
    }
  }
  c->box = b;
  c->nnodes = nn; c->nvel = 3 * nn; c->nel = ne; c->npres = 4 * ne; c->nbricks = nb;
  if ((rc = dev_alloc(c, (void **)&c->eta_bm, sizeof(double) * (size_t)nb * 27 * ptq2::BEL))) return rc;

  if ((rc = clear_err(c))) return rc;
  HIPCHK(c, ptq2::launch_check_mesh(b, c->coords, c->err, c->stream));
  int bits = 0;
  if ((rc = read_err(c, &bits))) return rc;
  if (bits & ptq2::ERRBIT_MESH) return fail(c, PTQ2_ERR_MESH, "non-positive Jacobian determinant in the mesh");
  c->mesh_ready = true;
  c->last.clear();
  return PTQ2_OK;
}

int ptq2_setup_mesh(ptq2_ctx *c, int mx, int my, int mz, const double *coords, const int32_t *elnidx,
                    const uint8_t *mask) {
  if (!c) return PTQ2_ERR_INVALID_ARG;
  if (mx < 1 || my < 1 || mz < 1) return fail(c, PTQ2_ERR_INVALID_ARG, "bad element counts");
  const int NX = 2 * mx + 1, NY = 2 * my + 1;
  if (elnidx) {
    int64_t e = 0;
    for (int ez = 0; ez < mz; ez++)
      for (int ey = 0; ey < my; ey++)
        for (int ex = 0; ex < mx; ex++, e++)
          for (int cc = 0; cc < 3; cc++)
            for (int bb = 0; bb < 3; bb++)
              for (int aa = 0; aa < 3; aa++) {
                int64_t want = (int64_t)(2 * ex + aa) + (int64_t)NX * ((int64_t)(2 * ey + bb) + (int64_t)NY * (2 * ez + cc));
                if ((int64_t)elnidx[27 * e + aa + 3 * bb + 9 * cc] != want)
                  return fail(c, PTQ2_ERR_UNSUPPORTED, "element %lld is not the structured Q2 map", (long long)e);
              }
  }
  return ptq2_setup_mesh_box(c, mx, my, mz, NX, NY, 2 * mz + 1, 0, 0, 0, coords, mask);
}

int ptq2_set_viscosity(ptq2_ctx *c, const double *eta, uint64_t state, unsigned flags) {
  int rc = check_ready(c, false);
  if (rc) return rc;
  if (!eta) return fail(c, PTQ2_ERR_INVALID_ARG, "null viscosity");
  if (flags & ~PTQ2_DEVICE) return fail(c, PTQ2_ERR_INVALID_ARG, "bad flags");
  if (c->eta_ready && c->has_state && c->state == state) return PTQ2_OK;  // cached (mf->state)
  if ((rc = set_device(c))) return rc;
  const double *src = eta;
  double *tmp = nullptr;
  const size_t bytes = sizeof(double) * 27 * (size_t)c->nel;
  if (!(flags & PTQ2_DEVICE)) {
    if ((rc = dev_alloc(c, (void **)&tmp, bytes))) return rc;
    hipError_t e = hipMemcpyAsync(tmp, eta, bytes, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) {
      dev_free(tmp);
      return fail(c, PTQ2_ERR_DEVICE, "viscosity upload: %s", hipGetErrorString(e));
    }
    src = tmp;
  }
  rc = clear_err(c);
  hipError_t e = hipSuccess;
  if (!rc) e = ptq2::launch_eta_to_bm(c->box, src, c->eta_bm, c->err, c->stream);
  int bits = 0;
  if (!rc && e == hipSuccess) rc = read_err(c, &bits);
  dev_free(tmp);
  if (rc) return rc;
  if (e != hipSuccess) return fail(c, PTQ2_ERR_DEVICE, "eta permutation: %s", hipGetErrorString(e));
  c->fac32_valid = false;
  c->facb32_valid = false;
  c->gen++;
  if (bits & ptq2::ERRBIT_VISCOSITY) {
    c->eta_ready = false;
    return fail(c, PTQ2_ERR_VISCOSITY, "eta <= 0 or non-finite at a quadrature point");
  }
  c->eta_ready = true;
  c->has_state = true;
  c->state = state;
  c->last.clear();
  return PTQ2_OK;
}

int ptq2_a11_apply(ptq2_ctx *c, const void *u, void *y, unsigned flags) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (!u || !y) return fail(c, PTQ2_ERR_INVALID_ARG, "null vector");
  if (flags & ~(PTQ2_DEVICE | PTQ2_FP32 | PTQ2_MATMULT | PTQ2_ADD | PTQ2_ONTHEFLY | PTQ2_STOREDGEO | PTQ2_MFMA))
    return fail(c, PTQ2_ERR_INVALID_ARG, "bad flags");
  if ((flags & PTQ2_MFMA) && !(flags & PTQ2_FP32)) return fail(c, PTQ2_ERR_INVALID_ARG, "PTQ2_MFMA needs PTQ2_FP32");
  if ((flags & PTQ2_MATMULT) && (flags & PTQ2_ADD)) return fail(c, PTQ2_ERR_INVALID_ARG, "MATMULT and ADD exclusive");
  if ((flags & PTQ2_ONTHEFLY) && (flags & PTQ2_STOREDGEO))
    return fail(c, PTQ2_ERR_INVALID_ARG, "ONTHEFLY and STOREDGEO exclusive");
  if ((rc = set_device(c))) return rc;
  const bool dev = flags & PTQ2_DEVICE, f32 = flags & PTQ2_FP32, mm = flags & PTQ2_MATMULT, add = flags & PTQ2_ADD;
  const bool otf = !(flags & PTQ2_STOREDGEO);
  const uint8_t *mask = mm ? c->mask : nullptr;
  const int64_t n = c->nvel;
  const size_t es = f32 ? sizeof(float) : sizeof(double);
  if (f32 && (rc = build_fac32(c))) return rc;
  if (!f32 && !otf && (rc = build_geo64(c))) return rc;
  const void *du = u;
  void *dy = y;
  if (!dev) {
    if ((rc = clear_err(c))) return rc;
    rc = f32 ? stage_in<float>(c, u, &c->su, &c->su_bytes, n, true) : stage_in<double>(c, u, &c->su, &c->su_bytes, n, true);
    if (rc) return rc;
    if ((rc = ensure(c, &c->sy, &c->sy_bytes, es * (size_t)n))) return rc;
    if (add) HIPCHK(c, hipMemcpyAsync(c->sy, y, es * (size_t)n, hipMemcpyHostToDevice, c->stream));
    du = c->su;
    dy = c->sy;
  }
  void *tb = ptq2_internal_facebuf(c, &rc);
  if (rc) return rc;
  if (f32) {
    if (flags & PTQ2_MFMA)
      HIPCHK(c, ptq2::launch_a11_fp32_mfma(c->box, (const float *)du, mask, c->fac32, (float *)dy, add, c->stream,
                                           (float *)tb));
    else
      HIPCHK(c, ptq2::launch_a11_fp32(c->box, (const float *)du, mask, c->fac32, (float *)dy, add, c->stream,
                                      (float *)tb));
    if (mm) HIPCHK(c, ptq2::launch_bc_rows<float>(c->ndir, c->dir_idx, (const float *)du, (float *)dy, 0, c->stream));
  } else {
    if (otf)
      HIPCHK(c, ptq2::launch_a11_fp64(c->box, c->coords, (const double *)du, mask, c->eta_bm, (double *)dy, add,
                                      c->stream, (double *)tb));
    else
      HIPCHK(c, ptq2::launch_a11_fp64g(c->box, (const double *)du, mask, c->geo64, c->eta_bm, (double *)dy, add,
                                       c->stream, (double *)tb));
    if (mm) HIPCHK(c, ptq2::launch_bc_rows<double>(c->ndir, c->dir_idx, (const double *)du, (double *)dy, 0, c->stream));
  }
  if (!dev) {
    HIPCHK(c, hipMemcpyAsync(y, c->sy, es * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    int bits = 0;
    if ((rc = read_err(c, &bits))) return rc;
    if (bits & ptq2::ERRBIT_NONFINITE) return fail(c, PTQ2_ERR_NONFINITE, "non-finite value in the input vector");
  }
  return PTQ2_OK;
}

int ptq2_a11_apply_noalloc(ptq2_ctx *c, const double *u, double *y, unsigned flags) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (!u || !y) return fail(c, PTQ2_ERR_INVALID_ARG, "null vector");
  if (!(flags & PTQ2_DEVICE)) return fail(c, PTQ2_ERR_INVALID_ARG, "PTQ2_DEVICE is required (host vectors need staging)");
  if (flags & ~(PTQ2_DEVICE | PTQ2_MATMULT | PTQ2_ADD | PTQ2_ONTHEFLY))
    return fail(c, PTQ2_ERR_INVALID_ARG, "bad flags (FP64, geometry from the coordinates only)");
  if ((flags & PTQ2_MATMULT) && (flags & PTQ2_ADD)) return fail(c, PTQ2_ERR_INVALID_ARG, "MATMULT and ADD exclusive");
  const size_t need = sizeof(double) * (size_t)c->nbricks * ptq2::FACEVALS;
  if (need > NOALLOC_SCRATCH_BYTES)
    return fail(c, PTQ2_ERR_SIZE, "mesh of %lld bricks needs %zu B of face slots, above the %zu B scratch",
                (long long)c->nbricks, need, NOALLOC_SCRATCH_BYTES);
  if ((rc = set_device(c))) return rc;
  double *scr = nullptr;
  HIPCHK(c, scratch_resolve(c->device, &scr));  // resolved by ptq2_create: a table lookup
  const bool mm = flags & PTQ2_MATMULT, add = flags & PTQ2_ADD;
  // every face slot the gather pass reads is written by the apply launch before it in stream
  // order, so the array needs no clearing between calls
  HIPCHK(c, ptq2::launch_a11_fp64(c->box, c->coords, u, mm ? c->mask : nullptr, c->eta_bm, y, add, c->stream, scr));
  if (mm) HIPCHK(c, ptq2::launch_bc_rows<double>(c->ndir, c->dir_idx, u, y, 0, c->stream));
  return PTQ2_OK;
}

}  // extern "C"

namespace {
int ctx_event(ptq2_ctx *c) {
  if (!c->ev) HIPCHK(c, hipEventCreateWithFlags(&c->ev, hipEventDisableTiming));
  return PTQ2_OK;
}

template <typename T>
int a11_batch(ptq2_ctx *const *cs, int n, const void *const *u, void *const *y, unsigned flags) {
  ptq2_ctx *c0 = cs[0];
  const bool mm = flags & PTQ2_MATMULT, add = flags & PTQ2_ADD, f32 = sizeof(T) == 4;
  int rc = set_device(c0);
  if (rc) return rc;
  ptq2::BatchArgs<T> a{};
  a.n = n;
  a.add = add ? 1 : 0;
  int64_t blocks = 0;
  for (int i = 0; i < n; i++) {
    ptq2_ctx *c = cs[i];
    rc = check_ready(c, true);
    if (rc) return rc == PTQ2_ERR_INVALID_ARG ? fail(c0, rc, "null context %d", i) : fail(c0, rc, "problem %d: %s", i, c->last.c_str());
    if (c->device != c0->device) return fail(c0, PTQ2_ERR_INVALID_ARG, "problem %d is on device %d, not %d", i, c->device, c0->device);
    if (!u[i] || !y[i]) return fail(c0, PTQ2_ERR_INVALID_ARG, "null vector of problem %d", i);
    for (int j = 0; j < i; j++)
      if (y[j] == y[i] || cs[j] == c) return fail(c0, PTQ2_ERR_INVALID_ARG, "problems %d and %d share a context or an output", j, i);
    if (f32 && (rc = build_fac32(c))) return fail(c0, rc, "problem %d: %s", i, c->last.c_str());
    void *fb = ptq2_internal_facebuf(c, &rc);
    if (rc) return fail(c0, rc, "problem %d: %s", i, c->last.c_str());
    a.start[i] = (int)blocks;
    a.box[i] = c->box;
    a.u[i] = (const T *)u[i];
    a.y[i] = (T *)y[i];
    a.face[i] = (T *)fb;
    a.mask[i] = mm ? c->mask : nullptr;
    a.coords[i] = c->coords;
    if constexpr (sizeof(T) == 4)
      a.fac[i] = c->fac32;
    else
      a.fac[i] = c->eta_bm;
    blocks += (c->nbricks + 7) / 8 * 8;
    if (blocks > INT32_MAX) return fail(c0, PTQ2_ERR_SIZE, "batch of more than 2^31 bricks");
  }
  a.start[n] = (int)blocks;
  // the launch runs on the first problem's stream, ordered after and before the work of every
  // other problem's stream
  if ((rc = ctx_event(c0))) return rc;
  for (int i = 1; i < n; i++) {
    ptq2_ctx *c = cs[i];
    if (c->stream == c0->stream) continue;
    if ((rc = ctx_event(c))) return fail(c0, rc, "problem %d: %s", i, c->last.c_str());
    HIPCHK(c0, hipEventRecord(c->ev, c->stream));
    HIPCHK(c0, hipStreamWaitEvent(c0->stream, c->ev, 0));
  }
  HIPCHK(c0, ptq2::launch_a11_batch<T>(a, c0->stream));
  if (mm)
    for (int i = 0; i < n; i++)
      HIPCHK(c0, ptq2::launch_bc_rows<T>(cs[i]->ndir, cs[i]->dir_idx, (const T *)u[i], (T *)y[i], 0, c0->stream));
  bool others = false;
  for (int i = 1; i < n; i++) others |= cs[i]->stream != c0->stream;
  if (others) {
    HIPCHK(c0, hipEventRecord(c0->ev, c0->stream));
    for (int i = 1; i < n; i++)
      if (cs[i]->stream != c0->stream) HIPCHK(c0, hipStreamWaitEvent(cs[i]->stream, c0->ev, 0));
  }
  return PTQ2_OK;
}
}  // namespace

extern "C" {

static_assert(ptq2::BATCH_MAX == PTQ2_BATCH_MAX, "batch bound of the header and the kernels");

int ptq2_a11_apply_batch(ptq2_ctx *const *ctxs, int n, const void *const *u, void *const *y, unsigned flags) {
  if (!ctxs || !ctxs[0]) return PTQ2_ERR_INVALID_ARG;
  ptq2_ctx *c0 = ctxs[0];
  if (n < 1 || n > ptq2::BATCH_MAX) return fail(c0, PTQ2_ERR_INVALID_ARG, "batch of %d problems (1..%d)", n, ptq2::BATCH_MAX);
  if (!u || !y) return fail(c0, PTQ2_ERR_INVALID_ARG, "null vector array");
  if (!(flags & PTQ2_DEVICE)) return fail(c0, PTQ2_ERR_INVALID_ARG, "PTQ2_DEVICE is required");
  if (flags & ~(PTQ2_DEVICE | PTQ2_FP32 | PTQ2_MATMULT | PTQ2_ADD)) return fail(c0, PTQ2_ERR_INVALID_ARG, "bad flags");
  if ((flags & PTQ2_MATMULT) && (flags & PTQ2_ADD)) return fail(c0, PTQ2_ERR_INVALID_ARG, "MATMULT and ADD exclusive");
  return (flags & PTQ2_FP32) ? a11_batch<float>(ctxs, n, u, y, flags) : a11_batch<double>(ctxs, n, u, y, flags);
}

int ptq2_a11_diag(ptq2_ctx *c, double *d, unsigned flags) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (!d) return fail(c, PTQ2_ERR_INVALID_ARG, "null vector");
  if (flags & ~(PTQ2_DEVICE | PTQ2_MATMULT | PTQ2_ADD)) return fail(c, PTQ2_ERR_INVALID_ARG, "bad flags");
  if ((flags & PTQ2_MATMULT) && (flags & PTQ2_ADD)) return fail(c, PTQ2_ERR_INVALID_ARG, "MATMULT and ADD exclusive");
  if ((rc = set_device(c))) return rc;
  const bool dev = flags & PTQ2_DEVICE, mm = flags & PTQ2_MATMULT, add = flags & PTQ2_ADD;
  const int64_t n = c->nvel;
  double *dd = d;
  if (!dev) {
    if ((rc = ensure(c, &c->sy, &c->sy_bytes, sizeof(double) * (size_t)n))) return rc;
    if (add) HIPCHK(c, hipMemcpyAsync(c->sy, d, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    dd = (double *)c->sy;
  }
  HIPCHK(c, ptq2::launch_diag(c->box, c->coords, c->eta_bm, dd, add, c->stream));
  if (mm) HIPCHK(c, ptq2::launch_bc_rows<double>(c->ndir, c->dir_idx, (const double *)nullptr, dd, 2, c->stream));
  if (!dev) {
    HIPCHK(c, hipMemcpyAsync(d, dd, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return PTQ2_OK;
}

int ptq2_a12_apply(ptq2_ctx *c, const double *p, double *yu, unsigned flags) {
  int rc = check_ready(c, false);
  if (rc) return rc;
  if (!p || !yu) return fail(c, PTQ2_ERR_INVALID_ARG, "null vector");
  if (flags & ~(PTQ2_DEVICE | PTQ2_MATMULT | PTQ2_ADD)) return fail(c, PTQ2_ERR_INVALID_ARG, "bad flags");
  if ((flags & PTQ2_MATMULT) && (flags & PTQ2_ADD)) return fail(c, PTQ2_ERR_INVALID_ARG, "MATMULT and ADD exclusive");
  if ((rc = set_device(c))) return rc;
  const bool dev = flags & PTQ2_DEVICE, mm = flags & PTQ2_MATMULT, add = flags & PTQ2_ADD;
  const double *dp = p;
  double *dy = yu;
  if (!dev) {
    if ((rc = clear_err(c))) return rc;
    if ((rc = stage_in<double>(c, p, (void **)&c->sp, &c->sp_bytes, c->npres, true))) return rc;
    if ((rc = ensure(c, &c->sy, &c->sy_bytes, sizeof(double) * (size_t)c->nvel))) return rc;
    if (add) HIPCHK(c, hipMemcpyAsync(c->sy, yu, sizeof(double) * (size_t)c->nvel, hipMemcpyHostToDevice, c->stream));
    dp = c->sp;
    dy = (double *)c->sy;
  }
  HIPCHK(c, ptq2::launch_a12(c->box, c->coords, dp, dy, add, c->stream));
  if (mm) HIPCHK(c, ptq2::launch_bc_rows<double>(c->ndir, c->dir_idx, (const double *)nullptr, dy, 1, c->stream));
  if (!dev) {
    HIPCHK(c, hipMemcpyAsync(yu, dy, sizeof(double) * (size_t)c->nvel, hipMemcpyDeviceToHost, c->stream));
    int bits = 0;
    if ((rc = read_err(c, &bits))) return rc;
    if (bits & ptq2::ERRBIT_NONFINITE) return fail(c, PTQ2_ERR_NONFINITE, "non-finite value in the input vector");
  }
  return PTQ2_OK;
}

int ptq2_body_force(ptq2_ctx *c, const double *fu, double *y, unsigned flags) {
  int rc = check_ready(c, false);
  if (rc) return rc;
  if (!fu || !y) return fail(c, PTQ2_ERR_INVALID_ARG, "null vector");
  if (flags & ~(PTQ2_DEVICE | PTQ2_MATMULT | PTQ2_ADD)) return fail(c, PTQ2_ERR_INVALID_ARG, "bad flags");
  if ((flags & PTQ2_MATMULT) && (flags & PTQ2_ADD)) return fail(c, PTQ2_ERR_INVALID_ARG, "MATMULT and ADD exclusive");
  if ((rc = set_device(c))) return rc;
  const bool dev = flags & PTQ2_DEVICE, mm = flags & PTQ2_MATMULT, add = flags & PTQ2_ADD;
  const double *df = fu;
  double *dy = y;
  if (!dev) {
    if ((rc = clear_err(c))) return rc;
    // the pressure staging buffer, grown to the 81 values per element of fu
    if ((rc = stage_in<double>(c, fu, (void **)&c->sp, &c->sp_bytes, 81 * c->nel, true))) return rc;
    if ((rc = ensure(c, &c->sy, &c->sy_bytes, sizeof(double) * (size_t)c->nvel))) return rc;
    if (add) HIPCHK(c, hipMemcpyAsync(c->sy, y, sizeof(double) * (size_t)c->nvel, hipMemcpyHostToDevice, c->stream));
    df = c->sp;
    dy = (double *)c->sy;
  }
  void *fb = ptq2_internal_facebuf(c, &rc);
  if (rc) return rc;
  HIPCHK(c, ptq2::launch_body_force(c->box, c->coords, df, dy, add, (double *)fb, c->stream));
  if (mm) HIPCHK(c, ptq2::launch_bc_rows<double>(c->ndir, c->dir_idx, (const double *)nullptr, dy, 1, c->stream));
  if (!dev) {
    HIPCHK(c, hipMemcpyAsync(y, dy, sizeof(double) * (size_t)c->nvel, hipMemcpyDeviceToHost, c->stream));
    int bits = 0;
    if ((rc = read_err(c, &bits))) return rc;
    if (bits & ptq2::ERRBIT_NONFINITE) return fail(c, PTQ2_ERR_NONFINITE, "non-finite value in the body force");
  }
  return PTQ2_OK;
}

int ptq2_a21_apply(ptq2_ctx *c, const double *u, double *yp, unsigned flags) {
  int rc = check_ready(c, false);
  if (rc) return rc;
  if (!u || !yp) return fail(c, PTQ2_ERR_INVALID_ARG, "null vector");
  if (flags & ~(PTQ2_DEVICE | PTQ2_MATMULT | PTQ2_ADD)) return fail(c, PTQ2_ERR_INVALID_ARG, "bad flags");
  if ((flags & PTQ2_MATMULT) && (flags & PTQ2_ADD)) return fail(c, PTQ2_ERR_INVALID_ARG, "MATMULT and ADD exclusive");
  if ((rc = set_device(c))) return rc;
  const bool dev = flags & PTQ2_DEVICE, mm = flags & PTQ2_MATMULT, add = flags & PTQ2_ADD;
  const double *du = u;
  double *dy = yp;
  if (!dev) {
    if ((rc = clear_err(c))) return rc;
    if ((rc = stage_in<double>(c, u, &c->su, &c->su_bytes, c->nvel, true))) return rc;
    if ((rc = ensure(c, (void **)&c->syp, &c->syp_bytes, sizeof(double) * (size_t)c->npres))) return rc;
    if (add) HIPCHK(c, hipMemcpyAsync(c->syp, yp, sizeof(double) * (size_t)c->npres, hipMemcpyHostToDevice, c->stream));
    du = (const double *)c->su;
    dy = c->syp;
  }
  HIPCHK(c, ptq2::launch_a21(c->box, c->coords, du, mm ? c->mask : nullptr, dy, add, c->stream));
  if (!dev) {
    HIPCHK(c, hipMemcpyAsync(yp, dy, sizeof(double) * (size_t)c->npres, hipMemcpyDeviceToHost, c->stream));
    int bits = 0;
    if ((rc = read_err(c, &bits))) return rc;
    if (bits & ptq2::ERRBIT_NONFINITE) return fail(c, PTQ2_ERR_NONFINITE, "non-finite value in the input vector");
  }
  return PTQ2_OK;
}

int ptq2_stokes_apply(ptq2_ctx *c, const void *u, const void *p, void *yu, void *yp, unsigned flags) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (!u || !p || !yu || !yp) return fail(c, PTQ2_ERR_INVALID_ARG, "null vector");
  if (flags & ~(PTQ2_DEVICE | PTQ2_FP32 | PTQ2_MATMULT | PTQ2_ADD | PTQ2_ONTHEFLY | PTQ2_STOREDGEO))
    return fail(c, PTQ2_ERR_INVALID_ARG, "bad flags");
  if ((flags & PTQ2_MATMULT) && (flags & PTQ2_ADD)) return fail(c, PTQ2_ERR_INVALID_ARG, "MATMULT and ADD exclusive");
  if ((flags & PTQ2_ONTHEFLY) && (flags & PTQ2_STOREDGEO))
    return fail(c, PTQ2_ERR_INVALID_ARG, "ONTHEFLY and STOREDGEO exclusive");
  if ((rc = set_device(c))) return rc;
  const bool dev = flags & PTQ2_DEVICE, f32 = flags & PTQ2_FP32, mm = flags & PTQ2_MATMULT, add = flags & PTQ2_ADD;
  const bool otf = !(flags & PTQ2_STOREDGEO);
  const uint8_t *mask = mm ? c->mask : nullptr;
  const size_t es = f32 ? sizeof(float) : sizeof(double);
  if (f32 && (rc = ptq2_internal_build_fac32(c, true))) return rc;
  if (!f32 && !otf && (rc = build_geo64(c))) return rc;
  const void *du = u, *dp = p;
  void *dyu = yu, *dyp = yp;
  if (!dev) {
    if ((rc = clear_err(c))) return rc;
    if (f32) {
      if ((rc = stage_in<float>(c, u, &c->su, &c->su_bytes, c->nvel, true))) return rc;
      if ((rc = stage_in<float>(c, p, (void **)&c->sp, &c->sp_bytes, c->npres, true))) return rc;
    } else {
      if ((rc = stage_in<double>(c, u, &c->su, &c->su_bytes, c->nvel, true))) return rc;
      if ((rc = stage_in<double>(c, p, (void **)&c->sp, &c->sp_bytes, c->npres, true))) return rc;
    }
    if ((rc = ensure(c, &c->sy, &c->sy_bytes, es * (size_t)c->nvel))) return rc;
    if ((rc = ensure(c, (void **)&c->syp, &c->syp_bytes, es * (size_t)c->npres))) return rc;
    if (add) {
      HIPCHK(c, hipMemcpyAsync(c->sy, yu, es * (size_t)c->nvel, hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipMemcpyAsync(c->syp, yp, es * (size_t)c->npres, hipMemcpyHostToDevice, c->stream));
    }
    du = c->su; dp = c->sp; dyu = c->sy; dyp = c->syp;
  }
  void *fb = ptq2_internal_facebuf(c, &rc);
  if (rc) return rc;
  if (f32) {
    // one fused pass (MFStokesWrapper_A_AVX's single element loop), FP32 factors
    HIPCHK(c, ptq2::launch_stokes_fp32(c->box, (const float *)du, (const float *)dp, mask, c->fac32, c->facb32,
                                       (float *)dyu, (float *)dyp, add, c->stream, (float *)fb));
    if (mm) HIPCHK(c, ptq2::launch_bc_rows<float>(c->ndir, c->dir_idx, (const float *)du, (float *)dyu, 0, c->stream));
  } else {
    if (otf) {
      // momentum A11 u + A12 p and continuity A21 u in one element pass (MFStokesWrapper_A_AVX)
      HIPCHK(c, ptq2::launch_stokes_fp64(c->box, c->coords, (const double *)du, (const double *)dp, mask, c->eta_bm,
                                         (double *)dyu, (double *)dyp, nullptr, add, c->stream, (double *)fb));
    } else {
      // resident geometry: the A11 kernel of that form, then A12 and A21
      HIPCHK(c, ptq2::launch_a11_fp64g(c->box, (const double *)du, mask, c->geo64, c->eta_bm, (double *)dyu, add,
                                       c->stream, (double *)fb));
      HIPCHK(c, ptq2::launch_a12(c->box, c->coords, (const double *)dp, (double *)dyu, true, c->stream));
      HIPCHK(c, ptq2::launch_a21(c->box, c->coords, (const double *)du, mask, (double *)dyp, add, c->stream));
    }
    if (mm)
      HIPCHK(c, ptq2::launch_bc_rows<double>(c->ndir, c->dir_idx, (const double *)du, (double *)dyu, 0, c->stream));
  }
  if (!dev) {
    HIPCHK(c, hipMemcpyAsync(yu, dyu, es * (size_t)c->nvel, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(yp, dyp, es * (size_t)c->npres, hipMemcpyDeviceToHost, c->stream));
    int bits = 0;
    if ((rc = read_err(c, &bits))) return rc;
    if (bits & ptq2::ERRBIT_NONFINITE) return fail(c, PTQ2_ERR_NONFINITE, "non-finite value in the input vector");
  }
  return PTQ2_OK;
}

int ptq2_stokes_residual(ptq2_ctx *c, const double *u, const double *p, const double *fu, const double *fp,
                         const double *phi, double *Fu, double *Fp, unsigned flags) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (!u || !p || !fu || !Fu || !Fp) return fail(c, PTQ2_ERR_INVALID_ARG, "null vector");
  if (flags & ~PTQ2_DEVICE) return fail(c, PTQ2_ERR_INVALID_ARG, "bad flags");
  if ((rc = set_device(c))) return rc;
  const bool dev = flags & PTQ2_DEVICE;
  const int64_t nv = c->nvel, np = c->npres, ne = c->nel;
  const double *du = u, *dp = p, *dfu = fu, *dfp = fp, *dphi = phi;
  double *dFu = Fu, *dFp = Fp;
  // host path: one staging buffer holds u, p, fu, fp and phi back to back
  double *stage = nullptr;
  struct Free {
    double *p;
    ~Free() { if (p) (void)hipFree(p); }
  } fr{nullptr};
  if (!dev) {
    const int64_t n = nv + np + 81 * ne + (fp ? 27 * ne : 0) + (phi ? nv : 0) + nv + np;
    if ((rc = dev_alloc(c, (void **)&stage, sizeof(double) * (size_t)n))) return rc;
    fr.p = stage;
    double *o = stage;
    auto put = [&](const double *h, int64_t k) -> double * {
      double *d = o;
      (void)hipMemcpyAsync(d, h, sizeof(double) * (size_t)k, hipMemcpyHostToDevice, c->stream);
      o += k;
      return d;
    };
    du = put(u, nv);
    dp = put(p, np);
    dfu = put(fu, 81 * ne);
    dfp = fp ? put(fp, 27 * ne) : nullptr;
    dphi = phi ? put(phi, nv) : nullptr;
    dFu = o;
    dFp = o + nv;
    HIPCHK(c, hipGetLastError());
    if ((rc = clear_err(c))) return rc;
    HIPCHK(c, ptq2::launch_scan_finite<double>(o - stage, stage, c->err, c->stream));
  }
  // u_loc: u with the Dirichlet values inserted (BCListInsertLocal, src/dmda_bcs.c:547-573)
  const double *ul = du;
  if (c->ndir) {
    if ((rc = ensure(c, (void **)&c->ures, &c->ures_bytes, sizeof(double) * (size_t)nv))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->ures, du, sizeof(double) * (size_t)nv, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, ptq2::launch_bc_rows<double>(c->ndir, c->dir_idx, dphi, c->ures, dphi ? 0 : 1, c->stream));
    ul = c->ures;
  }
  void *fb = ptq2_internal_facebuf(c, &rc);
  if (rc) return rc;
  // Fu = A11 u_loc + A12 p, Fp = A21 u_loc - int N^p fp (one element pass), then Fu -= Be
  HIPCHK(c, ptq2::launch_stokes_fp64(c->box, c->coords, ul, dp, nullptr, c->eta_bm, dFu, dFp, dfp, false, c->stream,
                                     (double *)fb));
  HIPCHK(c, ptq2::launch_body_force(c->box, c->coords, dfu, dFu, true, (double *)fb, c->stream, -1.0));
  // Dirichlet rows: Fu_D = u_D - phi_D (BCListResidualDirichlet, src/dmda_bcs.c:614-648)
  HIPCHK(c, ptq2::launch_bc_rows<double>(c->ndir, c->dir_idx, du, dFu, 3, c->stream, dphi));
  if (!dev) {
    HIPCHK(c, hipMemcpyAsync(Fu, dFu, sizeof(double) * (size_t)nv, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(Fp, dFp, sizeof(double) * (size_t)np, hipMemcpyDeviceToHost, c->stream));
    int bits = 0;
    if ((rc = read_err(c, &bits))) return rc;
    if (bits & ptq2::ERRBIT_NONFINITE) return fail(c, PTQ2_ERR_NONFINITE, "non-finite value in an input vector");
  }
  return PTQ2_OK;
}

}  // extern "C"
