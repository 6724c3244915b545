/*
 * ptq2.h -- C ABI of the MI355X matrix-free Q2-P1disc Stokes operator library
 * (libptq2.so).  Plain C types only: the PETSc-side adapter
 * (ptatin-gene_amd/adapter/stokes_operators_hip.c), the solver host code and the
 * tests bind to exactly these entry points.
 *
 * Mesh: one structured Q2 node box of NX x NY x NZ nodes holding mx x my x mz
 * owned elements whose first node sits at (i0,j0,k0) of the box; node
 * n = i + NX*(j + NY*k), interlaced velocity dofs 3n+l, element
 * e = ex + mx*(ey + my*ez), pressure dofs 4e+d, quadrature point q = a+3b+9c.
 * A single-rank DMDA is NX=2mx+1, (i0,j0,k0)=(0,0,0); a rank's ghosted local
 * box is the general case.  This is the numbering of DMDAGetElements_DA_Q2_3D
 * (src/dmda_element_q2p1.c:368-469) and DMDAGetElements_DA_P0MD_3D (:491-546).
 *
 * Every function returns PTQ2_OK (0) or a nonzero PTQ2_ERR_* code; no function
 * aborts the process.  A context is not thread safe; independent contexts are.
 */
#ifndef PTQ2_H
#define PTQ2_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ptq2_ctx ptq2_ctx;

enum {
  PTQ2_OK = 0,
  PTQ2_ERR_INVALID_ARG = 1, /* null pointer, bad flag or bad dimension */
  PTQ2_ERR_NOT_SETUP = 2,   /* mesh or viscosity missing (SETERRQ "MF operator not setup") */
  PTQ2_ERR_SIZE = 3,        /* size overflow or above the supported maximum */
  PTQ2_ERR_ALLOC = 4,       /* device or host allocation failed */
  PTQ2_ERR_DEVICE = 5,      /* HIP runtime error */
  PTQ2_ERR_NONFINITE = 6,   /* NaN or Inf in coordinates, viscosity or input vector */
  PTQ2_ERR_MESH = 7,        /* non-positive Jacobian determinant at a quadrature point */
  PTQ2_ERR_VISCOSITY = 8,   /* eta <= 0 at a quadrature point */
  PTQ2_ERR_UNSUPPORTED = 9  /* connectivity that is not a structured DMDA box */
};

/* flags of the apply entry points */
#define PTQ2_DEVICE 0x1u  /* vectors are device pointers; the call is stream ordered and
                             returns before the work completes (ptq2_sync to wait); device
                             input vectors are not scanned for NaN/Inf (the host path does
                             that, and ptq2_mg_solve checks its right-hand side's norm) */
#define PTQ2_FP32 0x2u    /* float vectors; A11 runs the FP32 smoother operator */
#define PTQ2_MATMULT 0x4u /* MatShell semantics: Dirichlet columns zeroed on input and
                             Dirichlet rows replaced on output */
#define PTQ2_ADD 0x8u     /* y += (element sum), the SpMVOp raw-kernel contract;
                             without it y = (element sum) */
#define PTQ2_ONTHEFLY 0x10u /* FP64 A11: recompute the element geometry from the coordinates
                               in every apply (the default; the flag only makes it explicit) */
#define PTQ2_STOREDGEO 0x20u /* FP64 A11: read resident per-point geometric factors instead
                               (1944 B per element, built on the first such apply after a
                               mesh set-up); slower than recomputing on MI355X */
#define PTQ2_MFMA 0x40u /* with PTQ2_FP32: the FP32 operator with the 81x27 reference gradient
                           applied as a dense GEMM on the matrix cores (same factors; FP32
                           rounding differs from the sum-factorized kernel's) */

/* lifecycle ------------------------------------------------------------- */
/* replaces MFA11CUDA_SetUp (include/stokes_operators.h:124-149) and the
   SpMVOp_SetUp hook of struct _p_MatA11MF (include/stokes_operators.h:48-64) */
int ptq2_create(ptq2_ctx **ctx, int device);
/* replaces MFA11CUDA_CleanUp / SpMVOp_Destroy */
int ptq2_destroy(ptq2_ctx **ctx);
const char *ptq2_strerror(int code);
/* message of the last failing call on this context ("" when none) */
const char *ptq2_last_error(const ptq2_ctx *ctx);
/* run subsequent work on the hipStream_t `hip_stream` (NULL is the null stream, as used by
   a caller's default stream); a new context runs on a stream of its own */
int ptq2_set_stream(ptq2_ctx *ctx, void *hip_stream);
void *ptq2_get_stream(const ptq2_ctx *ctx);
int ptq2_sync(ptq2_ctx *ctx);

/* mesh and coefficients ---------------------------------------------------- */
/* Uploads the coordinates (interlaced FP64, 3*NX*NY*NZ) once; they stay resident.
   elnidx (27*mx*my*mz, may be NULL) is checked against the structured map of the
   box; dirichlet_mask (3*NX*NY*NZ bytes, nonzero = Dirichlet dof, may be NULL) is
   the BCList of src/dmda_bcs.c.  Replaces the connectivity/coordinate uploads of
   CopyTo_A11_CUDA (src/stokes_operators_cuda.cu:325-431). */
int ptq2_setup_mesh(ptq2_ctx *ctx, int mx, int my, int mz, const double *coords,
                    const int32_t *elnidx, const uint8_t *dirichlet_mask);
int ptq2_setup_mesh_box(ptq2_ctx *ctx, int mx, int my, int mz, int NX, int NY, int NZ,
                        int i0, int j0, int k0, const double *coords,
                        const uint8_t *dirichlet_mask);
/* Quadrature-point viscosity eta_qp[27*e+q] (the .eta of QPntVolCoefStokes,
   include/QPntVolCoefStokes_def.h:41-46).  `state` mirrors mf->state
   (src/stokes_operators.c:1176-1177): an upload with the state already cached is
   skipped.  flags may hold PTQ2_DEVICE.  The FP32 operator's geometric factors
   (FP64 Jacobian, then rounded) are rebuilt on the next FP32 apply. */
int ptq2_set_viscosity(ptq2_ctx *ctx, const double *eta_qp, uint64_t state, unsigned flags);

/* operators ----------------------------------------------------------------- */
/* A11 viscous block.  Raw mode (no PTQ2_MATMULT) is the SpMVOp_MatMult contract of
   MFStokesWrapper_A11_* (include/stokes_operators_mf.h:36-70):
   ufield[] local velocity in, element contributions summed into Yu[].
   PTQ2_MATMULT is MatMult_MFStokes_A11 (src/stokes_operators.c:1156-1222). */
int ptq2_a11_apply(ptq2_ctx *ctx, const void *u, void *y, unsigned flags);
/* No counterpart in the reference.  ptq2_a11_apply for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
   the same operator and results, FP64 with the geometry recomputed from the coordinates; flags
   must hold PTQ2_DEVICE and may add PTQ2_MATMULT or PTQ2_ADD.  The face slots of the one-launch
   output path live in a library-owned static device array of 32 MiB per GPU instead of a
   per-context buffer, so a mesh above 5753 bricks of 4x4x4 elements (e.g. beyond 68^3 elements)
   is PTQ2_ERR_SIZE.  Calls that overlap on one GPU (any stream, any context) race on that array;
   the library never allocates in this call.  Stream ordered, like the other PTQ2_DEVICE calls. */
int ptq2_a11_apply_noalloc(ptq2_ctx *ctx, const double *u, double *y, unsigned flags);
/* ptq2_a11_apply of n independent problems (1..PTQ2_BATCH_MAX contexts on one device, each
   its own mesh and viscosity, meshes of any sizes) in ONE launch of the A11 kernel and one of
   the face pass: y[i] = A11_i u[i].  The reference applies one operator per MatMult
   (MatMult_MFStokes_A11, src/stokes_operators.c:1156-1222); the batch serves many small
   sinker problems (the 24^3-80^3 batch) whose single launches would each fill a fraction of
   the GPU.  flags must hold PTQ2_DEVICE and may add PTQ2_FP32 (float vectors, the smoother
   operator), PTQ2_MATMULT or PTQ2_ADD; FP64 recomputes the geometry from the coordinates.
   Results are bitwise those of ptq2_a11_apply on each problem.  The launch is ordered on
   ctxs[0]'s stream after the work already queued on every ctxs[i]'s stream, and every
   ctxs[i]'s stream waits for it.  Contexts and outputs must be distinct.  Device vectors are
   not scanned for NaN/Inf. */
#define PTQ2_BATCH_MAX 16
int ptq2_a11_apply_batch(ptq2_ctx *const *ctxs, int n, const void *const *u, void *const *y, unsigned flags);
/* diagonal of A11: MFStokesWrapper_diagA11 (src/stokes_operators_mf.c:71-135);
   with PTQ2_MATMULT, MatGetDiagonal_MFStokes_A11 (src/stokes_operators.c:1072-1113)
   which puts 1 in Dirichlet slots.  FP64 only. */
int ptq2_a11_diag(ptq2_ctx *ctx, double *d, unsigned flags);
/* gradient A12 (pressure 4E in, velocity out): MFStokesWrapper_A12_AVX
   (src/stokes_operators_avx.c:682-829); PTQ2_MATMULT is MatMult_MFStokes_A12
   (src/stokes_operators.c:1301-1370): Dirichlet rows of y set to 0. */
int ptq2_a12_apply(ptq2_ctx *ctx, const double *p, double *yu, unsigned flags);
/* body-force load: yu[3n+d] += sum_e sum_q w_q detJ_q N_n(x_q) fu[(27e+q)*3+d], the Be term of
   FormFunctionLocal_U (src/stokes_form_function.c:219-224) with fu the Fu[3] of
   QPntVolCoefStokes (include/QPntVolCoefStokes_def.h:41-46); 81E doubles in.  The
   momentum residual of the reference is A11 u + A12 p - this vector.  PTQ2_MATMULT: Dirichlet
   rows of yu set to 0 (the right-hand side of the constrained solve). */
int ptq2_body_force(ptq2_ctx *ctx, const double *fu, double *yu, unsigned flags);
/* divergence A21 (velocity in, pressure 4E out): MFStokesWrapper_A21
   (src/stokes_operators_mf.c:862-937); PTQ2_MATMULT is MatMult_MFStokes_A21
   (src/stokes_operators.c:1444-1509): Dirichlet columns of u zeroed. */
int ptq2_a21_apply(ptq2_ctx *ctx, const double *u, double *yp, unsigned flags);
/* full block [A11 A12; A21 0]: MFStokesWrapper_A_AVX (src/stokes_operators_avx.c:439-636);
   PTQ2_MATMULT is MatMult_MFStokes_A (src/stokes_operators.c:893-979).  With PTQ2_FP32
   all four vectors are float and the block runs in one fused pass on FP32-stored
   geometric factors (FP64 Jacobians rounded once per viscosity state); otherwise double. */
int ptq2_stokes_apply(ptq2_ctx *ctx, const void *u, const void *p, void *yu, void *yp, unsigned flags);

/* Stokes residual of FormFunction_Stokes on one rank: FormFunctionLocal_U
   (src/stokes_form_function.c:99-242) and FormFunctionLocal_P (:478-580).  With u_loc = u whose
   Dirichlet dofs carry the values phi (BCListInsertLocal; phi 3N, NULL = zero values):
     Fu = A11 u_loc + A12 p - Be(fu),    Fu_D = u_D - phi_D (BCListResidualDirichlet),
     Fp = A21 u_loc - int N^p fp,
   fu the Fu[3] of QPntVolCoefStokes per quadrature point (81 per element, [e][q][d]) and fp its
   Fp (27 per element, may be NULL).  Both outputs are overwritten.  The momentum and continuity
   parts run in one element pass (the FP64 full-block kernel) plus the body-force load.  flags:
   PTQ2_DEVICE (stream ordered; device pointers are not scanned for NaN/Inf). */
int ptq2_stokes_residual(ptq2_ctx *ctx, const double *u, const double *p, const double *fu, const double *fp,
                         const double *phi, double *Fu, double *Fp, unsigned flags);
/* material points -> quadrature points: SwarmUpdateGaussPropertiesLocalL2Projection_Q1_MPntPStokes
   (src/material_point_utils.c:905-955, with its body :663-830; -coefficient_projection_type 1).
   npoints points, each with local coordinates xi[3p..3p+2] in its element wil[p]
   (0 <= wil < n_elements: the xi/wil fields of MPntStd) and values eta_p[p], rho_p[p] (the
   eta/rho of MPntPStokes; rho_p may be NULL).  Writes eta_qp and rho_qp (27 per element, q =
   a+3b+9c; rho_qp is NULL exactly when rho_p is), as the reference writes .eta/.rho of
   QPntVolCoefStokes.  Every node sums the points of its adjacent elements in a fixed order
   (ascending element, then point index), so the result is bitwise reproducible; it equals the
   reference's point-order sums to rounding.  flags: PTQ2_DEVICE (every array a device pointer).
   Synchronous (device scratch is allocated and freed inside).  Needs a single-rank box
   (ptq2_setup_mesh, or ptq2_setup_mesh_box without a ghost layer); a box with a ghost layer is
   PTQ2_ERR_UNSUPPORTED.  An element index out of range is PTQ2_ERR_INVALID_ARG, a non-finite
   coordinate or value PTQ2_ERR_NONFINITE.  Pass eta_qp to ptq2_set_viscosity to use it. */
int ptq2_project_points_q1(ptq2_ctx *ctx, int64_t npoints, const double *xi, const int32_t *wil,
                           const double *eta_p, const double *rho_p, double *eta_qp, double *rho_qp,
                           unsigned flags);
/* introspection (sizes, for callers that allocate device vectors) */
int ptq2_sizes(const ptq2_ctx *ctx, int64_t *n_velocity_dofs, int64_t *n_pressure_dofs,
               int64_t *n_elements);

#ifdef __cplusplus
}
#endif
#endif /* PTQ2_H */
