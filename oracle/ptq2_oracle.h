/*
 * ptq2_oracle.h -- CPU FP64 restatement of pTatin3D's matrix-free Q2-P1disc
 * Stokes operators.  TEST INFRASTRUCTURE ONLY: this library is the checker
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * It is never linked into, loaded by or called from the product library.
 *
 * Conventions (all follow the reference):
 *  - Structured DMDA Q2 mesh, mx*my*mz elements, nodes NX=2mx+1 (etc.),
 *    node n = i + NX*(j + NY*k), coordinates interlaced coords[3n+d]
 *    (src/ptatin3d_stokes.c:137,153; src/dmda_element_q2p1.c:368-469).
 *  - Element e = ex + mx*(ey + my*ez); local node / QP index i + 3j + 9k
 *    with i along x fastest (src/element_utils_q2.c:110-118,288-298).
 *  - Velocity dof = 3*node + component; pressure dof = 4*e + d
 *    (src/dmda_element_q2p1.c:491-546).
 *  - eta_qp[27*e + q] holds the quadrature-point viscosity
 *    (include/QPntVolCoefStokes_def.h:41-46, .eta field only).
 *
 * Element kernels ADD into their outputs ("raw kernel" semantics of the
 * SpMVOp seam, include/stokes_operators.h:48-64).  The *_matmult helpers add
 * the BC pre/post steps of src/stokes_operators.c:893-1509.
 */
#ifndef PTQ2_ORACLE_H
#define PTQ2_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* mesh */
void ptq2o_mesh_uniform(int mx, int my, int mz, double x0, double x1, double y0, double y1,
                        double z0, double z1, double *coords);
void ptq2o_mesh_bump_ymax(int mx, int my, int mz, double amp, double lambda, double *coords);
void ptq2o_elnidx(int mx, int my, int mz, int32_t *elnidx);
/* BC masks: type 0 = FreeSlipFreeSurface, 1 = FreeSlip, 2 = NoSlip */
void ptq2o_bc_mask(int mx, int my, int mz, int type, uint8_t *mask);
/* reference test vectors (src/tests/test_stokes_operators.c:48-130) */
void ptq2o_test_vector_u(int mx, int my, int mz, const double *coords, double *u);
void ptq2o_test_vector_p(int mx, int my, int mz, int legacy, double *p);

/* element operators, raw kernel semantics (y += ...) */
void ptq2o_a11_tensor(int mx, int my, int mz, const double *coords, const double *eta,
                      const double *u, double *y);
void ptq2o_a11_tensor_omp(int mx, int my, int mz, const double *coords, const double *eta,
                          const double *u, double *y, int nthreads);
void ptq2o_a11_ref(int mx, int my, int mz, const double *coords, const double *eta,
                   const double *u, double *y);
void ptq2o_a12(int mx, int my, int mz, const double *coords, const double *p, double *yu);
void ptq2o_a21(int mx, int my, int mz, const double *coords, const double *u, double *yp);
void ptq2o_diag_a11(int mx, int my, int mz, const double *coords, const double *eta, double *d);
void ptq2o_stokes_a(int mx, int my, int mz, const double *coords, const double *eta,
                    const double *u, const double *p, double *yu, double *yp);

/* shells with Dirichlet handling (one rank, local == global) */
void ptq2o_matmult_a11(int mx, int my, int mz, const double *coords, const double *eta,
                       const uint8_t *mask, const double *x, double *y);
void ptq2o_matmult_a12(int mx, int my, int mz, const double *coords, const uint8_t *mask,
                       const double *p, double *yu);
void ptq2o_matmult_a21(int mx, int my, int mz, const double *coords, const uint8_t *mask,
                       const double *u, double *yp);
void ptq2o_matgetdiag_a11(int mx, int my, int mz, const double *coords, const double *eta,
                          const uint8_t *mask, double *d);
void ptq2o_matmult_a(int mx, int my, int mz, const double *coords, const double *eta,
                     const uint8_t *mask, const double *u, const double *p, double *yu,
                     double *yp);

/* dense 81x81 element matrix of A11 at one element (brute-force check) */
void ptq2o_a11_element_matrix(const double *elcoords /*81*/, const double *eta /*27*/,
                              double *Ke /*81*81, row = 3*node + comp*/);

/* FormFunction_Stokes residual (momentum Fu 3N, continuity Fp 4E); fp and phi may be NULL */
void ptq2o_residual(int mx, int my, int mz, const double *coords, const double *eta, const uint8_t *mask,
                    const double *phi, const double *u, const double *p, const double *fu, const double *fp,
                    double *Fu, double *Fp);
/* y += body-force load of fu (81 per element) */
void ptq2o_body_force(int mx, int my, int mz, const double *coords, const double *fu, double *y);

/* material points: the zero-perturbation lattice layout (nx*ny*nz per element; xi 3 per point,
   wil = element, xp = physical coordinate) and the Q1 projection of point viscosity/density onto
   the quadrature points (rho_p / rho_qp may be NULL) */
void ptq2o_marker_lattice(int mx, int my, int mz, const double *coords, int nx, int ny, int nz, double *xi,
                          int32_t *wil, double *xp);
void ptq2o_project_q1(int mx, int my, int mz, int64_t npoints, const double *xi, const int32_t *wil,
                      const double *eta_p, const double *rho_p, double *eta_qp, double *rho_qp);

#ifdef __cplusplus
}
#endif
#endif
