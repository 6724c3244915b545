// ptq2_ctx.h -- the context behind the opaque ptq2_ctx handle (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "ptq2_device.h"

struct ptq2_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  bool mesh_ready = false, eta_ready = false;
  ptq2::Box box{};
  int64_t nnodes = 0, nvel = 0, nel = 0, npres = 0, nbricks = 0;
  double *coords = nullptr;
  uint8_t *mask = nullptr;
  int64_t *dir_idx = nullptr;
  int64_t ndir = 0;
  double *eta_bm = nullptr;
  float *fac32 = nullptr;
  bool fac32_valid = false;
  double *geo64 = nullptr;  // FP64 per-point geometry sqrt(w/det) adj(J) (mesh only)
  bool geo64_valid = false;
  float *facb32 = nullptr;  // FP32 full-block factors (w det / eta scale, pressure basis)
  bool facb32_valid = false;
  bool has_state = false;
  uint64_t state = 0;
  // bumped by every mesh set-up and every viscosity upload: a multigrid hierarchy built on this
  // context records it and refuses to run once it has changed (its captured graph holds the
  // device buffers of that mesh, and its coarse levels were derived from that viscosity)
  uint64_t gen = 0;
  int *err = nullptr;
  // output path of the A11 applies: all bricks in one launch, each storing the nodes it owns
  // and its upper-face values into its slot of facebuf (FACEVALS doubles per brick), then a
  // pass adds the face slots into their owners (ptq2::launch_a11_*).  One launch keeps the
  // memory phases of some bricks overlapped with the arithmetic of others, which 8 coloured
  // launches of a few waves per SIMD do not.
  void *facebuf = nullptr;
  // u with the Dirichlet values inserted, for the residual (allocated on first use)
  double *ures = nullptr;
  size_t ures_bytes = 0;
  // stream ordering of the batched apply (created on first use)
  hipEvent_t ev = nullptr;
  // staging for host-pointer calls
  void *su = nullptr, *sy = nullptr;
  size_t su_bytes = 0, sy_bytes = 0;
  double *sp = nullptr, *syp = nullptr;
  size_t sp_bytes = 0, syp_bytes = 0;
  std::string last;
};


// (re)build the FP32 geometric factors of the smoother operator if stale; with block,
// also those of the FP32 full block
int ptq2_internal_build_fac32(ptq2_ctx *c, bool block = false);
// the tile buffer of the uncoloured output path if this mesh uses it (allocated on first
// use), else nullptr; *rc receives an allocation error
void *ptq2_internal_facebuf(ptq2_ctx *c, int *rc);
