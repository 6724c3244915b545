/*
 * ptq2_solver.h -- geometric-multigrid preconditioned Krylov solve of the viscous
 * block A11 u = f on one GPU (libptq2.so).
 *
 * This replaces the PETSc configuration that pTatin3D builds around its A11 shells:
 * PCMG over a DMDA hierarchy coarsened by 2 (pTatin3dStokesKSPConfigureFSGMG,
 * src/ptatin_driver_linear_ts.c:820-883), Chebyshev smoothing with Jacobi
 * (KSPSolve_ChebychevRN_lag_norm, src/ksp_chebyrn.c:278-447; eigen-estimate transform
 * of KSPChebychevSetEstimateEigenvalues, :178-214), trilinear DMDA interpolation on the
 * Q2 node lattice with injected coordinates (src/dmda_project_coords.c:39-71), an
 * assembled FP64 coarse level (src/ptatin_driver_linear_ts.c:630-633) and FGMRES outside.
 *
 * Precision: smoothing runs in FP32 (the FP32 A11 operator with FP32 vectors) or in
 * FP64; the Krylov residual, the interlevel FP64 data and the coarse solve are FP64.
 * The whole V-cycle is replayed from a captured hipGraph.
 */
#ifndef PTQ2_SOLVER_H
#define PTQ2_SOLVER_H
#include <stdint.h>

#include "ptq2.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ptq2_mg ptq2_mg;

typedef struct {
  int nlevels;          /* levels including the fine one (>= 2); each coarsens by 2 */
  int smooth_its;       /* Chebyshev iterations per pre/post smoothing */
  int fp32_smoother;    /* 1: FP32 smoothing on every non-coarse level, 0: FP64 */
  double emin_frac;     /* emin = emin_frac * emax_estimate (tform[1]) */
  double emax_frac;     /* emax = emax_frac * emax_estimate (tform[3]) */
  int esteig_its;       /* power iterations for the estimate of lambda_max(D^-1 A) */
  int use_graph;        /* 1: replay the V-cycle from a captured hipGraph */
  int coarse_max_dofs;  /* refuse a dense coarse level larger than this */
  int galerkin_coarse;  /* 1: the coarsest operator is R A P of the level above it (assembled
                           by probing), 0: rediscretised on the coarsest mesh */
  int coarse_eta_mean;  /* PTQ2_ETA_MEAN_*: how a coarse quadrature point combines the fine
                           quadrature viscosities of the child elements it touches (mean or
                           maximum) */
  int fp32_basis;       /* 1: FGMRES keeps its finished Arnoldi vectors in FP32 (candidates,
                           Z vectors, reductions and the true residual stay FP64): half the
                           Gram-Schmidt traffic; restarts recompute the FP64 residual.  0
                           (default): FP64 basis */
  /* Two-level correction for stiff inclusions (no counterpart in the reference, whose
     velocity solve is PCMG + FGMRES only; for high viscosity contrasts, where the rigid
     motions of the inclusions stall the Krylov solve).  When the quadrature viscosity spans
     at least rigid_min_contrast, an inclusion is a connected set of elements (sharing a node)
     holding a viscosity above sqrt(min * max); the 6 rigid motions of each of the
     rigid_max_inclusions largest form W, and the preconditioner becomes
     z = W y + M^-1 (r - A W y), y = (W^T A W)^-1 W^T r (A-DEF2 around the V-cycle M^-1).
     Set-up cost: 6 fine applies per inclusion.  Memory: A W on the nodes it reaches. */
  int rigid_deflation;      /* 1: on, 0: off (default) */
  double rigid_min_contrast; /* default 100 (>= 1) */
  int rigid_max_inclusions;  /* default 128 (1..128) */
  /* 1: each Chebyshev step is the apply kernel with the Jacobi scaling and the three-term
     update fused into its store (every owned node that no neighbour's face slot adds to is
     updated there), plus a pass over the remaining lower-face nodes; results are bitwise those
     of 0 (default): the apply storing A x, then one update pass over all dofs */
  int fused_smoother;
  /* 1 (default): at set-up, time each kernel of every level's smoother apply
     (PTQ2_MG_KERNEL_*: one wave per brick of 4x4x4 elements; three waves per brick, each
     running a slab of quadrature rows; with FP32 smoothing, the matrix-core form) with fast
     or general tile addressing where every tile is full, and keep the fastest
     (ptq2_mg_tuning reports the choice).  The kernels sum in different orders, so two
     hierarchies whose choices differ round differently in the last bits.  0: one wave per
     brick and the mesh's own tile addressing on every level.  2: the three-wave kernel on
     every level, 3: the matrix-core form on every level (FP32 smoothing), 4: the one-wave
     kernel, and R A P through the level above on a Galerkin level below a matrix-free one,
     untimed (for tests and comparisons).  A Galerkin level below a matrix-free one is timed
     as its stencil and as R A P (prolongation, the level above's matrix-free apply,
     restriction), which streams far fewer bytes than the stencil on large levels. */
  int autotune;
  /* k >= 0 (default 0): the k levels above the coarsest carry the Galerkin operator
     Pi R A P Pi + (I - Pi) of the level above them instead of the rediscretised matrix-free
     operator (the reference's Galerkin coarse levels, src/ptatin_driver_linear_ts.c:64-385),
     assembled on the device by coloured probing as 5x5x5 stencils of 3x3 blocks
     (1125 values per node, FP64 plus an FP32 copy with FP32 smoothing).  0 <= k <=
     nlevels - 2; level 0 stays matrix-free.  For high viscosity contrasts. */
  int galerkin_levels;
} ptq2_mg_opts;

/* PTQ2_ETA_MEAN_Q1PROJ: the reference's coefficient hierarchy
   (SwarmUpdateGaussPropertiesLocalL2Projection_Q1_MPntPStokes_Hierarchy,
   src/material_point_utils.c:1325-1524, projection type 0): a nodal field on the fine Q2 node
   lattice (Q1 weights on the element corners, 1 elsewhere), restricted lattice by lattice with
   the scaled trilinear restriction, Q1-interpolated to each coarse level's quadrature points */
enum {
  PTQ2_ETA_MEAN_ARITHMETIC = 0,
  PTQ2_ETA_MEAN_GEOMETRIC = 1,
  PTQ2_ETA_MEAN_HARMONIC = 2,
  PTQ2_ETA_MEAN_MAX = 3,
  PTQ2_ETA_MEAN_Q1PROJ = 4
};

/* fill opts with the defaults: Chebyshev(4), FP32 smoothing, (0.2, 1.1) transform,
   20 power iterations, graph replay, coarse dense limit 20000 dofs, Galerkin coarsest
   operator, arithmetic coarse viscosity */
void ptq2_mg_default_opts(ptq2_mg_opts *opts);

/* Ownership: from ptq2_mg_create to ptq2_mg_destroy the hierarchy owns the fine context.  Its
   captured V-cycle holds the fine context's device buffers (coordinates, mask, viscosity,
   factors, face slots) and its coarse levels were derived from the fine viscosity, so a new
   ptq2_setup_mesh* or ptq2_set_viscosity upload on the fine context makes ptq2_mg_vcycle and
   ptq2_mg_solve return PTQ2_ERR_NOT_SETUP (rebuild the hierarchy).  The hierarchy moves the
   fine context onto its own stream (ptq2_mg_set_stream moves both); operator calls on the fine
   context must stay on that stream, since they share its face-slot buffer with the V-cycle.
   Device-pointer calls do not scan their vectors element by element: ptq2_mg_solve rejects a
   right-hand side whose norm is not finite (PTQ2_ERR_NONFINITE).

   Build the hierarchy under a set-up context (mesh with BC mask + viscosity): coarse
   meshes by coordinate injection, BC masks by injection, quadrature-point viscosity
   by averaging over the 8 children of each coarse element, Jacobi diagonals,
   eigen-estimates, the dense FP64 inverse of the coarsest operator, with rigid_deflation the
   stiff-inclusion correction, and the graph. */
int ptq2_mg_create(ptq2_mg **mg, ptq2_ctx *fine, const ptq2_mg_opts *opts);
int ptq2_mg_destroy(ptq2_mg **mg);
int ptq2_mg_levels(const ptq2_mg *mg, int *nlevels, int *m_per_level /* [nlevels] x,y,z triples */);

/* the stiff inclusions of the two-level correction: their number, the rigid-motion columns
   kept in (W^T A W)^-1, and, if node_label is not NULL, the inclusion of every fine node
   ((2mx+1)(2my+1)(2mz+1) entries, -1 outside all; 0 inclusions when the option is off) */
int ptq2_mg_rigid_info(const ptq2_mg *mg, int *ninclusions, int *ncolumns, int32_t *node_label);

/* the Chebyshev interval [emin, emax] of every smoothed level (nlevels-1 entries each) */
int ptq2_mg_eigs(const ptq2_mg *mg, double *emin, double *emax);
/* Kernels of the smoother apply (ptq2_mg_tuning) */
enum {
  PTQ2_MG_KERNEL_BRICK = 0,       /* one wave per brick of 4x4x4 elements */
  PTQ2_MG_KERNEL_SPLIT = 1,       /* three waves per brick, one slab of quadrature rows each */
  PTQ2_MG_KERNEL_MFMA = 2,        /* the FP32 operator's matrix-core form */
  PTQ2_MG_KERNEL_STENCIL = 3,     /* Galerkin level: its assembled 5x5x5 block stencil */
  PTQ2_MG_KERNEL_GALERKIN_MF = 4  /* Galerkin level: R A P through the matrix-free level above */
};
/* Launch shape the autotuner chose for the smoother apply of each level l < nlevels - 1:
   kernel (PTQ2_MG_KERNEL_*), tile addressing (1: fast full-tile path, 0: general), and the
   measured ms per apply of the chosen and of the default shape (both 0 when autotune was off
   or forced).  Any output may be NULL. */
int ptq2_mg_tuning(const ptq2_mg *mg, int *kernel, int *fast_tiles, double *ms, double *ms_default);
/* Per-level time of the V-cycle: reps V-cycles z = M^-1 r launched kernel by kernel (not from
   the graph) with events between the levels; ms[l] = the mean time of level l's own work
   (smoothing, residual and transfers; on the coarsest level its solve).  Synchronous. */
int ptq2_mg_level_times(ptq2_mg *mg, const double *r, double *z, int reps, double *ms);

/* order all subsequent solver work (every level and the captured V-cycle) on
   `hip_stream` (NULL = the null stream) */
int ptq2_mg_set_stream(ptq2_mg *mg, void *hip_stream);

/* one V-cycle z = M^-1 r on device FP64 vectors (stream ordered) */
int ptq2_mg_vcycle(ptq2_mg *mg, const double *r, double *z);

/* FGMRES(restart) with the V-cycle as right preconditioner on device FP64 vectors:
   solves MatMult_MFStokes_A11 x = b from x = 0 until ||b - A x|| <= rtol ||b||.
   1 <= restart <= 256; the solver keeps 2*restart+1 device vectors of the fine size
   (high-contrast problems with many stiff inclusions need long restarts: their
   near-rigid modes form a cluster of small eigenvalues).  Returns the iteration count
   and the final true residual norm. */
int ptq2_mg_solve(ptq2_mg *mg, const double *b, double *x, double rtol, int max_it, int restart,
                  int *iterations, double *final_rnorm, double *rnorm0);

/* residual history of the last solve (one norm per iteration, up to cap entries) */
/* 1 if the last ptq2_mg_solve stopped at the rounding floor: a restart cycle whose residual
   estimate reached rtol left a true residual (b - A x in FP64) above 1.5 rtol that had
   fallen by less than 10% since the cycle began, so rtol lies below what the FP64 residual of
   this problem resolves (stiff
   inclusions: their element sums cancel only to ~eps times the inclusion viscosity).  The
   solve then returns with that true residual in final_rnorm instead of running to max_it. */
int ptq2_mg_stagnated(const ptq2_mg *mg);
int ptq2_mg_history(const ptq2_mg *mg, double *hist, int cap, int *n);

/* message of the last failing call on this hierarchy ("" when none) */
const char *ptq2_mg_last_error(const ptq2_mg *mg);

#ifdef __cplusplus
}
#endif
#endif /* PTQ2_SOLVER_H */
