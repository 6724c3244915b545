/*
 * ptq2_oracle.c -- FP64 CPU restatement of pTatin3D's matrix-free Q2-P1disc
 * Stokes operator path.  TEST INFRASTRUCTURE ONLY (see ptq2_oracle.h).
 *
 * Two independent restatements of A11 are kept on purpose:
 *  - ptq2o_a11_tensor follows the sum-factorized algorithm of
 *    src/stokes_operators_tensor.c:40-286 (TensorContractNEV, JacobianInvertNEV,
 *    QuadratureAction) element by element;
 *  - ptq2o_a11_ref follows the per-quadrature-point form of
 *    src/stokes_operators_mf.c:237-351 with the pointwise kernel
 *    MatMultMF_Stokes_MixedFEM3d_B11 (src/stokes_q2p1_mf_operators_def.c:151-257)
 *    and the geometry of P3D_evaluate_geometry_elementQ2
 *    (src/element_utils_q2.c:400-475).
 * A12 follows MFStokesWrapper_A12_AVX (src/stokes_operators_avx.c:682-829), A21
 * follows MFStokesWrapper_A21 (src/stokes_operators_mf.c:862-937), the diagonal
 * follows MFStokesWrapper_diagA11 (src/stokes_operators_mf.c:71-135) and the
 * full block follows MFStokesWrapper_A_AVX (src/stokes_operators_avx.c:439-636).
 *
 * Parity pins (tests/test_oracle_*.py): reference goldens of
 * tests/integration/operator/compare_mf_asm_8x9x11/expected and friends,
 * brute-force 81x81 element matrices, rigid-body null space, exact energies,
 * symmetry, A21 == A12^T, and the two A11 restatements against each other.
 */
#include "ptq2_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NQP 27

/* 1D Q2 Lagrange basis on nodes {-1,0,1} evaluated at the 3-point Gauss rule:
   B[q][n] = l_n(x_q), D[q][n] = l_n'(x_q), weights w1 (src/stokes_operators_tensor.c:193-206,
   which uses PetscDTGaussQuadrature(3,-1,1); the closed form is used here). */
static void basis1d(double B[3][3], double D[3][3], double w1[3], double x1[3]) {
  const double s = sqrt(0.6);
  x1[0] = -s; x1[1] = 0.0; x1[2] = s;
  w1[0] = 5.0 / 9.0; w1[1] = 8.0 / 9.0; w1[2] = 5.0 / 9.0;
  for (int i = 0; i < 3; i++) {
    double x = x1[i];
    B[i][0] = 0.5 * (x * x - x);
    B[i][1] = 1.0 - x * x;
    B[i][2] = 0.5 * (x * x + x);
    D[i][0] = x - 0.5;
    D[i][1] = -2.0 * x;
    D[i][2] = x + 0.5;
  }
}

static inline int64_t nodeid(int NX, int NY, int i, int j, int k) {
  return (int64_t)i + (int64_t)NX * ((int64_t)j + (int64_t)NY * (int64_t)k);
}

/* ------------------------------------------------------------------ mesh */

/* DMDASetUniformCoordinates on the Q2 node lattice: x_i = x0 + i*(x1-x0)/(NX-1) */
void ptq2o_mesh_uniform(int mx, int my, int mz, double x0, double x1, double y0, double y1,
                        double z0, double z1, double *coords) {
  int NX = 2 * mx + 1, NY = 2 * my + 1, NZ = 2 * mz + 1;
  double hx = (x1 - x0) / (NX - 1), hy = (y1 - y0) / (NY - 1), hz = (z1 - z0) / (NZ - 1);
  for (int k = 0; k < NZ; k++)
    for (int j = 0; j < NY; j++)
      for (int i = 0; i < NX; i++) {
        int64_t n = nodeid(NX, NY, i, j, k);
        coords[3 * n + 0] = x0 + i * hx;
        coords[3 * n + 1] = y0 + j * hy;
        coords[3 * n + 2] = z0 + k * hz;
      }
}

/* MeshDeformation_GaussianBump_YMAX (src/mesh_deformation.c:48-125): reset to
   [-1,1]^3, lift the column heights by amp*exp(lambda*(x^2+z^2))+1 with constant
   spacing in j, then rescale back to the original bounding box. */
void ptq2o_mesh_bump_ymax(int mx, int my, int mz, double amp, double lambda, double *coords) {
  int NX = 2 * mx + 1, NY = 2 * my + 1, NZ = 2 * mz + 1;
  int64_t N = (int64_t)NX * NY * NZ;
  double gmin[3] = {1e300, 1e300, 1e300}, gmax[3] = {-1e300, -1e300, -1e300};
  for (int64_t n = 0; n < N; n++)
    for (int d = 0; d < 3; d++) {
      double v = coords[3 * n + d];
      if (v < gmin[d]) gmin[d] = v;
      if (v > gmax[d]) gmax[d] = v;
    }
  ptq2o_mesh_uniform(mx, my, mz, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, coords);
  for (int k = 0; k < NZ; k++)
    for (int j = 0; j < NY; j++)
      for (int i = 0; i < NX; i++) {
        int64_t n = nodeid(NX, NY, i, j, k);
        double xn = coords[3 * n + 0], zn = coords[3 * n + 2];
        double yh = amp * exp(lambda * (xn * xn + zn * zn)) + 1.0;
        double dy = (yh + 1.0) / (double)(NY - 1);
        coords[3 * n + 1] = -1.0 + j * dy;
      }
  for (int64_t n = 0; n < N; n++)
    for (int d = 0; d < 3; d++)
      coords[3 * n + d] = (gmax[d] - gmin[d]) * (coords[3 * n + d] + 1.0) / 2.0 + gmin[d];
}

/* DMDAGetElements_DA_Q2_3D, one rank (src/dmda_element_q2p1.c:368-469) */
void ptq2o_elnidx(int mx, int my, int mz, int32_t *elnidx) {
  int NX = 2 * mx + 1, NY = 2 * my + 1;
  int64_t e = 0;
  for (int ek = 0; ek < mz; ek++)
    for (int ej = 0; ej < my; ej++)
      for (int ei = 0; ei < mx; ei++, e++)
        for (int c = 0; c < 3; c++)
          for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++)
              elnidx[27 * e + a + 3 * b + 9 * c] =
                  (int32_t)nodeid(NX, NY, 2 * ei + a, 2 * ej + b, 2 * ek + c);
}

/* viscous_sinker boundary conditions (src/models/viscous_sinker/model_viscous_sinker_ops.c:192-270):
   face traversals set the listed velocity component to a Dirichlet zero. */
void ptq2o_bc_mask(int mx, int my, int mz, int type, uint8_t *mask) {
  int NX = 2 * mx + 1, NY = 2 * my + 1, NZ = 2 * mz + 1;
  int64_t N = (int64_t)NX * NY * NZ;
  memset(mask, 0, (size_t)(3 * N));
  for (int k = 0; k < NZ; k++)
    for (int j = 0; j < NY; j++)
      for (int i = 0; i < NX; i++) {
        int64_t n = nodeid(NX, NY, i, j, k);
        int imin = (i == 0), imax = (i == NX - 1), jmin = (j == 0), jmax = (j == NY - 1),
            kmin = (k == 0), kmax = (k == NZ - 1);
        if (type == 0) { /* FreeSlipFreeSurface */
          if (imin || imax) mask[3 * n + 0] = 1;
          if (jmin) mask[3 * n + 1] = 1;
          if (kmin || kmax) mask[3 * n + 2] = 1;
        } else if (type == 1) { /* FreeSlip */
          if (imin || imax) mask[3 * n + 0] = 1;
          if (jmin || jmax) mask[3 * n + 1] = 1;
          if (kmin || kmax) mask[3 * n + 2] = 1;
        } else { /* NoSlip */
          if (imin || imax || jmin || jmax || kmin || kmax)
            mask[3 * n + 0] = mask[3 * n + 1] = mask[3 * n + 2] = 1;
        }
      }
}

/* _GenerateTestVector: u_l(node) = X + Y + Z + l + 3 */
void ptq2o_test_vector_u(int mx, int my, int mz, const double *coords, double *u) {
  int64_t N = (int64_t)(2 * mx + 1) * (2 * my + 1) * (2 * mz + 1);
  for (int64_t n = 0; n < N; n++) {
    double s = coords[3 * n] + coords[3 * n + 1] + coords[3 * n + 2];
    for (int l = 0; l < 3; l++) u[3 * n + l] = s + (double)l + 3.0;
  }
}

/* _GenerateTestVectorDAP: p_d(e) = (-1+i*hx) + (-1+j*hy) + (-1+k*hz) + d + 3 with
   hx = 2/M (src/tests/test_stokes_operators.c:114-117).  The reference's expected files
   were produced with hx = 2/(M-1) (legacy = 1): that spacing reproduces every A12
   golden, including the NaN of the 1x1x3 run (2/0). */
void ptq2o_test_vector_p(int mx, int my, int mz, int legacy, double *p) {
  double dx = 2.0 / (mx - legacy), dy = 2.0 / (my - legacy), dz = 2.0 / (mz - legacy);
  int64_t e = 0;
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++, e++) {
        double f = (-1.0 + i * dx) + (-1.0 + j * dy) + (-1.0 + k * dz);
        for (int d = 0; d < 4; d++) p[4 * e + d] = f + (double)d + 3.0;
      }
}

/* ------------------------------------------------------------ tensor A11 */

/* y[a*9+b*3+c] += sum T[a][k] S[b][j] R[c][i] x[k*9+j*3+i]; transpose mode uses
   R^T,S^T,T^T (TensorContractNEV, src/stokes_operators_tensor.c:40-92). */
static void contract3(const double Rf[3][3], const double Sf[3][3], const double Tf[3][3], int transpose,
                      const double *x, double *y) {
  double R[3][3], S[3][3], T[3][3], u[27], v[27];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      R[a][b] = transpose ? Rf[b][a] : Rf[a][b];
      S[a][b] = transpose ? Sf[b][a] : Sf[a][b];
      T[a][b] = transpose ? Tf[b][a] : Tf[a][b];
    }
  memset(u, 0, sizeof u);
  memset(v, 0, sizeof v);
  for (int kj = 0; kj < 9; kj++)
    for (int i = 0; i < 3; i++)
      for (int c = 0; c < 3; c++) u[kj * 3 + c] += R[c][i] * x[kj * 3 + i];
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int c = 0; c < 3; c++)
        for (int b = 0; b < 3; b++) v[(k * 3 + b) * 3 + c] += S[b][j] * u[(k * 3 + j) * 3 + c];
  for (int k = 0; k < 3; k++)
    for (int a = 0; a < 3; a++)
      for (int ji = 0; ji < 9; ji++) y[a * 9 + ji] += T[a][k] * v[k * 9 + ji];
}

/* JacobianInvertNEV (src/stokes_operators_tensor.c:94-125): dx[j][k][q] = dx_k/dxi_j on
   entry, (J^-1)[j][k] on exit; det is not sign-checked (as in the reference). */
static void jacobian_invert(double dx[3][3][NQP], double det[NQP]) {
  for (int q = 0; q < NQP; q++) {
    double a[3][3];
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) a[j][k] = dx[j][k][q];
    double b0 = (a[1][1] * a[2][2] - a[2][1] * a[1][2]);
    double b3 = -(a[1][0] * a[2][2] - a[2][0] * a[1][2]);
    double b6 = (a[1][0] * a[2][1] - a[2][0] * a[1][1]);
    double dt = a[0][0] * b0 + a[0][1] * b3 + a[0][2] * b6;
    double id = 1.0 / dt;
    dx[0][0][q] = id * b0;
    dx[0][1][q] = -id * (a[0][1] * a[2][2] - a[2][1] * a[0][2]);
    dx[0][2][q] = id * (a[0][1] * a[1][2] - a[1][1] * a[0][2]);
    dx[1][0][q] = id * b3;
    dx[1][1][q] = id * (a[0][0] * a[2][2] - a[2][0] * a[0][2]);
    dx[1][2][q] = -id * (a[0][0] * a[1][2] - a[1][0] * a[0][2]);
    dx[2][0][q] = id * b6;
    dx[2][1][q] = -id * (a[0][0] * a[2][1] - a[2][0] * a[0][1]);
    dx[2][2][q] = id * (a[0][0] * a[1][1] - a[1][0] * a[0][1]);
    det[q] = dt;
  }
}

static void gather_element(int mx, int my, int e, const double *f, double el[3][27]) {
  int NX = 2 * mx + 1, NY = 2 * my + 1;
  int ei = e % mx, ej = (e / mx) % my, ek = e / (mx * my);
  for (int c = 0; c < 3; c++)
    for (int b = 0; b < 3; b++)
      for (int a = 0; a < 3; a++) {
        int64_t n = nodeid(NX, NY, 2 * ei + a, 2 * ej + b, 2 * ek + c);
        for (int l = 0; l < 3; l++) el[l][a + 3 * b + 9 * c] = f[3 * n + l];
      }
}

static void scatter_element(int mx, int my, int e, const double el[3][27], double *f) {
  int NX = 2 * mx + 1, NY = 2 * my + 1;
  int ei = e % mx, ej = (e / mx) % my, ek = e / (mx * my);
  for (int c = 0; c < 3; c++)
    for (int b = 0; b < 3; b++)
      for (int a = 0; a < 3; a++) {
        int64_t n = nodeid(NX, NY, 2 * ei + a, 2 * ej + b, 2 * ek + c);
        for (int l = 0; l < 3; l++) f[3 * n + l] += el[l][a + 3 * b + 9 * c];
      }
}

/* geometry of one element from its 27 node coordinates (tensor form) */
static void element_geometry(const double elx[3][27], double dx[3][3][NQP], double det[NQP]) {
  double B[3][3], D[3][3], w1[3], x1[3];
  basis1d(B, D, w1, x1);
  memset(dx, 0, sizeof(double) * 3 * 3 * NQP);
  for (int l = 0; l < 3; l++) {
    contract3(D, B, B, 0, elx[l], dx[0][l]);
    contract3(B, D, B, 0, elx[l], dx[1][l]);
    contract3(B, B, D, 0, elx[l], dx[2][l]);
  }
  jacobian_invert(dx, det);
}

static void a11_element_tensor(const double elx[3][27], const double *eta, const double elu[3][27],
                               double elv[3][27]) {
  double B[3][3], D[3][3], w1[3], x1[3], w[NQP];
  double dx[3][3][NQP], det[NQP], du[3][3][NQP], dv[3][3][NQP];
  basis1d(B, D, w1, x1);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) w[(i * 3 + j) * 3 + k] = w1[i] * w1[j] * w1[k];
  element_geometry(elx, dx, det);
  memset(du, 0, sizeof du);
  for (int l = 0; l < 3; l++) {
    contract3(D, B, B, 0, elu[l], du[0][l]);
    contract3(B, D, B, 0, elu[l], du[1][l]);
    contract3(B, B, D, 0, elu[l], du[2][l]);
  }
  /* QuadratureAction (src/stokes_operators_tensor.c:127-179) */
  for (int q = 0; q < NQP; q++) {
    double dux[3][3], Du[6], Dv[6], dvx[3][3];
    for (int l = 0; l < 3; l++)
      for (int k = 0; k < 3; k++)
        dux[k][l] = du[0][l][q] * dx[k][0][q] + du[1][l][q] * dx[k][1][q] + du[2][l][q] * dx[k][2][q];
    Du[0] = dux[0][0];
    Du[1] = dux[1][1];
    Du[2] = dux[2][2];
    Du[3] = 0.5 * (dux[0][1] + dux[1][0]);
    Du[4] = 0.5 * (dux[0][2] + dux[2][0]);
    Du[5] = 0.5 * (dux[1][2] + dux[2][1]);
    for (int k = 0; k < 6; k++) Dv[k] = 2.0 * eta[q] * Du[k];
    dvx[0][0] = Dv[0]; dvx[0][1] = Dv[3]; dvx[0][2] = Dv[4];
    dvx[1][0] = Dv[3]; dvx[1][1] = Dv[1]; dvx[1][2] = Dv[5];
    dvx[2][0] = Dv[4]; dvx[2][1] = Dv[5]; dvx[2][2] = Dv[2];
    for (int l = 0; l < 3; l++)
      for (int k = 0; k < 3; k++)
        dv[k][l][q] = w[q] * det[q] *
                      (dvx[0][l] * dx[0][k][q] + dvx[1][l] * dx[1][k][q] + dvx[2][l] * dx[2][k][q]);
  }
  memset(elv, 0, sizeof(double) * 81);
  for (int l = 0; l < 3; l++) {
    contract3(D, B, B, 1, dv[0][l], elv[l]);
    contract3(B, D, B, 1, dv[1][l], elv[l]);
    contract3(B, B, D, 1, dv[2][l], elv[l]);
  }
}

void ptq2o_a11_tensor(int mx, int my, int mz, const double *coords, const double *eta,
                      const double *u, double *y) {
  int64_t nel = (int64_t)mx * my * mz;
  for (int64_t e = 0; e < nel; e++) {
    double elx[3][27], elu[3][27], elv[3][27];
    gather_element(mx, my, (int)e, coords, elx);
    gather_element(mx, my, (int)e, u, elu);
    a11_element_tensor(elx, eta + 27 * e, elu, elv);
    scatter_element(mx, my, (int)e, elv, y);
  }
}

/* Multi-threaded CPU baseline of the same algorithm.  The reference serialises its
   scatter with `omp critical` (src/stokes_operators_avx.c:289-291); here the scatter is
   made race free by processing the elements in 8 colours of (ex%2,ey%2,ez%2), which
   never share a node, so every colour is an ordinary parallel loop. */
void ptq2o_a11_tensor_omp(int mx, int my, int mz, const double *coords, const double *eta,
                          const double *u, double *y, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (int color = 0; color < 8; color++) {
    int cx = color & 1, cy = (color >> 1) & 1, cz = (color >> 2) & 1;
    int nx = (mx - cx + 1) / 2, ny = (my - cy + 1) / 2, nz = (mz - cz + 1) / 2;
    int64_t ncol = (int64_t)nx * ny * nz;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t t = 0; t < ncol; t++) {
      int ex = cx + 2 * (int)(t % nx), ey = cy + 2 * (int)((t / nx) % ny), ez = cz + 2 * (int)(t / ((int64_t)nx * ny));
      int e = ex + mx * (ey + my * ez);
      double elx[3][27], elu[3][27], elv[3][27];
      gather_element(mx, my, e, coords, elx);
      gather_element(mx, my, e, u, elu);
      a11_element_tensor(elx, eta + 27 * (int64_t)e, elu, elv);
      scatter_element(mx, my, e, elv, y);
    }
  }
}

/* --------------------------------------------------- per-QP ("ref") forms */

/* P3D_prepare_elementQ2_3x3 with the closed-form Gauss rule */
static void prepare_q2_3x3(double W[NQP], double XI[NQP][3], double NI[NQP][27], double GNI[NQP][3][27]) {
  double B[3][3], D[3][3], w1[3], x1[3];
  basis1d(B, D, w1, x1);
  int cnt = 0;
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++, cnt++) {
        XI[cnt][0] = x1[i]; XI[cnt][1] = x1[j]; XI[cnt][2] = x1[k];
        W[cnt] = w1[i] * w1[j] * w1[k];
        int n = 0;
        for (int c = 0; c < 3; c++)
          for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++, n++) {
              NI[cnt][n] = B[i][a] * B[j][b] * B[k][c];
              GNI[cnt][0][n] = D[i][a] * B[j][b] * B[k][c];
              GNI[cnt][1][n] = B[i][a] * D[j][b] * B[k][c];
              GNI[cnt][2][n] = B[i][a] * B[j][b] * D[k][c];
            }
      }
}

/* P3D_evaluate_geometry_elementQ2 (src/element_utils_q2.c:400-475) */
static void geometry_qp(const double GNI[3][27], const double elx[3][27], double *detJ, double dN[3][27]) {
  double J[3][3] = {{0}}, iJ[3][3];
  for (int n = 0; n < 27; n++)
    for (int d = 0; d < 3; d++)
      for (int l = 0; l < 3; l++) J[d][l] += GNI[d][n] * elx[l][n];
  double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double id = 1.0 / det;
  iJ[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  iJ[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) * id;
  iJ[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  iJ[1][0] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]) * id;
  iJ[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  iJ[1][2] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) * id;
  iJ[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  iJ[2][1] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) * id;
  iJ[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  for (int n = 0; n < 27; n++)
    for (int m = 0; m < 3; m++) dN[m][n] = iJ[m][0] * GNI[0][n] + iJ[m][1] * GNI[1][n] + iJ[m][2] * GNI[2][n];
  *detJ = det;
}

void ptq2o_a11_ref(int mx, int my, int mz, const double *coords, const double *eta,
                   const double *u, double *y) {
  static double W[NQP], XI[NQP][3], NI[NQP][27], GNI[NQP][3][27];
  prepare_q2_3x3(W, XI, NI, GNI);
  int64_t nel = (int64_t)mx * my * mz;
  for (int64_t e = 0; e < nel; e++) {
    double elx[3][27], elu[3][27], Ye[3][27];
    gather_element(mx, my, (int)e, coords, elx);
    gather_element(mx, my, (int)e, u, elu);
    memset(Ye, 0, sizeof Ye);
    for (int q = 0; q < NQP; q++) {
      double detJ, dN[3][27];
      geometry_qp(GNI[q], elx, &detJ, dN);
      double fac = W[q] * detJ, et = eta[27 * e + q];
      /* strain rate (e_xx,e_yy,e_zz,2e_xy,2e_xz,2e_yz) and tau = eta*(...) as in
         MatMultMF_Stokes_MixedFEM3d_B11 */
      double g[3][3] = {{0}};
      for (int n = 0; n < 27; n++)
        for (int l = 0; l < 3; l++)
          for (int m = 0; m < 3; m++) g[l][m] += elu[l][n] * dN[m][n];
      double sr[6] = {g[0][0], g[1][1], g[2][2], g[0][1] + g[1][0], g[0][2] + g[2][0], g[1][2] + g[2][1]};
      double tau[6] = {2 * et * sr[0], 2 * et * sr[1], 2 * et * sr[2], et * sr[3], et * sr[4], et * sr[5]};
      for (int n = 0; n < 27; n++) {
        Ye[0][n] += fac * (dN[0][n] * tau[0] + dN[1][n] * tau[3] + dN[2][n] * tau[4]);
        Ye[1][n] += fac * (dN[0][n] * tau[3] + dN[1][n] * tau[1] + dN[2][n] * tau[5]);
        Ye[2][n] += fac * (dN[0][n] * tau[4] + dN[1][n] * tau[5] + dN[2][n] * tau[2]);
      }
    }
    scatter_element(mx, my, (int)e, Ye, y);
  }
}

/* 81x81 element matrix K[3n+l][3m+k] = sum_q w detJ (eps-form), the brute-force form of
   src/stokes_assembly.c:219-269 (B^T D B with D = diag(2eta,2eta,2eta,eta,eta,eta)). */
void ptq2o_a11_element_matrix(const double *elcoords, const double *eta, double *Ke) {
  static double W[NQP], XI[NQP][3], NI[NQP][27], GNI[NQP][3][27];
  prepare_q2_3x3(W, XI, NI, GNI);
  double elx[3][27];
  for (int n = 0; n < 27; n++)
    for (int l = 0; l < 3; l++) elx[l][n] = elcoords[3 * n + l];
  memset(Ke, 0, sizeof(double) * 81 * 81);
  for (int q = 0; q < NQP; q++) {
    double detJ, dN[3][27];
    geometry_qp(GNI[q], elx, &detJ, dN);
    double fac = W[q] * detJ, et = eta[q];
    double Bm[6][81];
    memset(Bm, 0, sizeof Bm);
    for (int n = 0; n < 27; n++) {
      Bm[0][3 * n + 0] = dN[0][n];
      Bm[1][3 * n + 1] = dN[1][n];
      Bm[2][3 * n + 2] = dN[2][n];
      Bm[3][3 * n + 0] = dN[1][n]; Bm[3][3 * n + 1] = dN[0][n];
      Bm[4][3 * n + 0] = dN[2][n]; Bm[4][3 * n + 2] = dN[0][n];
      Bm[5][3 * n + 1] = dN[2][n]; Bm[5][3 * n + 2] = dN[1][n];
    }
    double Dd[6] = {2 * et, 2 * et, 2 * et, et, et, et};
    for (int r = 0; r < 81; r++)
      for (int c = 0; c < 81; c++) {
        double s = 0;
        for (int k = 0; k < 6; k++) s += Bm[k][r] * Dd[k] * Bm[k][c];
        Ke[r * 81 + c] += fac * s;
      }
  }
}

/* P3D_ConstructNi_P1GRel_3D (src/element_utils_q2.c:183-224) */
static void pressure_basis(const double xi[3], const double elx[3][27], double Np[4]) {
  double B1[3][3];
  for (int d = 0; d < 3; d++) {
    double x = xi[d];
    B1[d][0] = 0.5 * x * (x - 1.0);
    B1[d][1] = (1.0 + x) * (1.0 - x);
    B1[d][2] = 0.5 * (1.0 + x) * x;
  }
  double xg[3] = {0, 0, 0}, avg[3] = {0, 0, 0};
  int n = 0;
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++, n++) {
        double Ni = B1[0][i] * B1[1][j] * B1[2][k];
        for (int d = 0; d < 3; d++) {
          xg[d] += Ni * elx[d][n];
          avg[d] += elx[d][n];
        }
      }
  for (int d = 0; d < 3; d++) avg[d] *= (1.0 / 27.0);
  double L[3] = {elx[0][14] - elx[0][12], elx[1][16] - elx[1][10], elx[2][22] - elx[2][4]};
  Np[0] = 1.0;
  for (int d = 0; d < 3; d++) Np[1 + d] = (xg[d] - avg[d]) / L[d];
}

/* A12, tensor form of MFStokesWrapper_A12_AVX + QuadratureAction_A12_AVX */
void ptq2o_a12(int mx, int my, int mz, const double *coords, const double *p, double *yu) {
  double B[3][3], D[3][3], w1[3], x1[3], w[NQP], xi[NQP][3];
  basis1d(B, D, w1, x1);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) w[(i * 3 + j) * 3 + k] = w1[i] * w1[j] * w1[k];
  int q = 0;
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++, q++) { xi[q][0] = x1[i]; xi[q][1] = x1[j]; xi[q][2] = x1[k]; }
  int64_t nel = (int64_t)mx * my * mz;
  for (int64_t e = 0; e < nel; e++) {
    double elx[3][27], elv[3][27], dx[3][3][NQP], det[NQP], dv[3][3][NQP];
    gather_element(mx, my, (int)e, coords, elx);
    element_geometry(elx, dx, det);
    for (q = 0; q < NQP; q++) {
      double Np[4];
      pressure_basis(xi[q], elx, Np);
      double dp = 0;
      for (int i = 0; i < 4; i++) dp += Np[i] * p[4 * e + i];
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) dv[k][l][q] = w[q] * det[q] * (-dp) * dx[l][k][q];
    }
    memset(elv, 0, sizeof elv);
    for (int l = 0; l < 3; l++) {
      contract3(D, B, B, 1, dv[0][l], elv[l]);
      contract3(B, D, B, 1, dv[1][l], elv[l]);
      contract3(B, B, D, 1, dv[2][l], elv[l]);
    }
    scatter_element(mx, my, (int)e, elv, yu);
  }
}

/* A21, per-QP form of MFStokesWrapper_A21 + MatMultMF_Stokes_MixedFEM3d_A21 */
void ptq2o_a21(int mx, int my, int mz, const double *coords, const double *u, double *yp) {
  static double W[NQP], XI[NQP][3], NI[NQP][27], GNI[NQP][3][27];
  prepare_q2_3x3(W, XI, NI, GNI);
  int64_t nel = (int64_t)mx * my * mz;
  for (int64_t e = 0; e < nel; e++) {
    double elx[3][27], elu[3][27], Ye[4] = {0, 0, 0, 0};
    gather_element(mx, my, (int)e, coords, elx);
    gather_element(mx, my, (int)e, u, elu);
    for (int q = 0; q < NQP; q++) {
      double detJ, dN[3][27], Np[4];
      pressure_basis(XI[q], elx, Np);
      geometry_qp(GNI[q], elx, &detJ, dN);
      double fac = W[q] * detJ, div = 0;
      for (int n = 0; n < 27; n++) div += elu[0][n] * dN[0][n] + elu[1][n] * dN[1][n] + elu[2][n] * dN[2][n];
      for (int i = 0; i < 4; i++) Ye[i] += -fac * Np[i] * div;
    }
    for (int i = 0; i < 4; i++) yp[4 * e + i] += Ye[i];
  }
}

/* diagonal of A11, MFStokesWrapper_diagA11 + MatMultMF_Stokes_MixedFEM3d_diagB11 */
void ptq2o_diag_a11(int mx, int my, int mz, const double *coords, const double *eta, double *d) {
  static double W[NQP], XI[NQP][3], NI[NQP][27], GNI[NQP][3][27];
  prepare_q2_3x3(W, XI, NI, GNI);
  int64_t nel = (int64_t)mx * my * mz;
  for (int64_t e = 0; e < nel; e++) {
    double elx[3][27], Ye[3][27];
    gather_element(mx, my, (int)e, coords, elx);
    memset(Ye, 0, sizeof Ye);
    for (int q = 0; q < NQP; q++) {
      double detJ, dN[3][27];
      geometry_qp(GNI[q], elx, &detJ, dN);
      double fac = W[q] * detJ, et = eta[27 * e + q];
      for (int n = 0; n < 27; n++) {
        double ax = dN[0][n] * dN[0][n], ay = dN[1][n] * dN[1][n], az = dN[2][n] * dN[2][n];
        Ye[0][n] += fac * (et * ay + et * az + 2.0 * et * ax);
        Ye[1][n] += fac * (et * ax + et * az + 2.0 * et * ay);
        Ye[2][n] += fac * (et * ax + et * ay + 2.0 * et * az);
      }
    }
    scatter_element(mx, my, (int)e, Ye, d);
  }
}

/* full [A11 A12; A21 0] in one element pass (MFStokesWrapper_A_AVX + QuadratureAction_A_AVX) */
void ptq2o_stokes_a(int mx, int my, int mz, const double *coords, const double *eta,
                    const double *u, const double *p, double *yu, double *yp) {
  double B[3][3], D[3][3], w1[3], x1[3], w[NQP], xi[NQP][3];
  basis1d(B, D, w1, x1);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) w[(i * 3 + j) * 3 + k] = w1[i] * w1[j] * w1[k];
  int q = 0;
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++, q++) { xi[q][0] = x1[i]; xi[q][1] = x1[j]; xi[q][2] = x1[k]; }
  int64_t nel = (int64_t)mx * my * mz;
  for (int64_t e = 0; e < nel; e++) {
    double elx[3][27], elu[3][27], elv[3][27], dx[3][3][NQP], det[NQP], du[3][3][NQP], dv[3][3][NQP];
    double Np[NQP][4], dp[NQP], elq[4] = {0, 0, 0, 0};
    gather_element(mx, my, (int)e, coords, elx);
    gather_element(mx, my, (int)e, u, elu);
    for (q = 0; q < NQP; q++) pressure_basis(xi[q], elx, Np[q]);
    element_geometry(elx, dx, det);
    memset(du, 0, sizeof du);
    for (int l = 0; l < 3; l++) {
      contract3(D, B, B, 0, elu[l], du[0][l]);
      contract3(B, D, B, 0, elu[l], du[1][l]);
      contract3(B, B, D, 0, elu[l], du[2][l]);
    }
    for (q = 0; q < NQP; q++) {
      dp[q] = 0;
      for (int i = 0; i < 4; i++) dp[q] += Np[q][i] * p[4 * e + i];
    }
    for (q = 0; q < NQP; q++) {
      double dux[3][3], Du[6], Dv[6], dvx[3][3];
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++)
          dux[k][l] = du[0][l][q] * dx[k][0][q] + du[1][l][q] * dx[k][1][q] + du[2][l][q] * dx[k][2][q];
      Du[0] = dux[0][0]; Du[1] = dux[1][1]; Du[2] = dux[2][2];
      Du[3] = 0.5 * (dux[0][1] + dux[1][0]);
      Du[4] = 0.5 * (dux[0][2] + dux[2][0]);
      Du[5] = 0.5 * (dux[1][2] + dux[2][1]);
      for (int k = 0; k < 6; k++) Dv[k] = 2.0 * eta[27 * e + q] * Du[k];
      for (int k = 0; k < 3; k++) Dv[k] -= dp[q];
      dvx[0][0] = Dv[0]; dvx[0][1] = Dv[3]; dvx[0][2] = Dv[4];
      dvx[1][0] = Dv[3]; dvx[1][1] = Dv[1]; dvx[1][2] = Dv[5];
      dvx[2][0] = Dv[4]; dvx[2][1] = Dv[5]; dvx[2][2] = Dv[2];
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++)
          dv[k][l][q] = w[q] * det[q] * (dvx[0][l] * dx[0][k][q] + dvx[1][l] * dx[1][k][q] + dvx[2][l] * dx[2][k][q]);
      /* continuity: -div(u) tested with N^p */
      double gx = dux[0][0] + dux[1][1] + dux[2][2];
      for (int i = 0; i < 4; i++) elq[i] += w[q] * (Np[q][i] * (-gx)) * det[q];
    }
    memset(elv, 0, sizeof elv);
    for (int l = 0; l < 3; l++) {
      contract3(D, B, B, 1, dv[0][l], elv[l]);
      contract3(B, D, B, 1, dv[1][l], elv[l]);
      contract3(B, B, D, 1, dv[2][l], elv[l]);
    }
    scatter_element(mx, my, (int)e, elv, yu);
    for (int i = 0; i < 4; i++) yp[4 * e + i] += elq[i];
  }
}

/* ---------------------------------------------------------------- shells */

static int64_t nvel(int mx, int my, int mz) { return 3 * (int64_t)(2 * mx + 1) * (2 * my + 1) * (2 * mz + 1); }

/* MatMult_MFStokes_A11 (src/stokes_operators.c:1156-1222): zero Dirichlet columns,
   apply, y_k = x_k on Dirichlet rows (BCListInsertDirichlet_MatMult, scale = 1). */
void ptq2o_matmult_a11(int mx, int my, int mz, const double *coords, const double *eta,
                       const uint8_t *mask, const double *x, double *y) {
  int64_t n = nvel(mx, my, mz);
  double *xl = (double *)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) { xl[i] = (mask && mask[i]) ? 0.0 : x[i]; y[i] = 0.0; }
  ptq2o_a11_tensor(mx, my, mz, coords, eta, xl, y);
  if (mask)
    for (int64_t i = 0; i < n; i++)
      if (mask[i]) y[i] = x[i];
  free(xl);
}

/* MatMult_MFStokes_A12 (src/stokes_operators.c:1301-1370): Dirichlet rows of y set to 0 */
void ptq2o_matmult_a12(int mx, int my, int mz, const double *coords, const uint8_t *mask,
                       const double *p, double *yu) {
  int64_t n = nvel(mx, my, mz);
  memset(yu, 0, sizeof(double) * n);
  ptq2o_a12(mx, my, mz, coords, p, yu);
  if (mask)
    for (int64_t i = 0; i < n; i++)
      if (mask[i]) yu[i] = 0.0;
}

/* MatMult_MFStokes_A21 (src/stokes_operators.c:1444-1509): Dirichlet columns of u zeroed */
void ptq2o_matmult_a21(int mx, int my, int mz, const double *coords, const uint8_t *mask,
                       const double *u, double *yp) {
  int64_t n = nvel(mx, my, mz), np = 4 * (int64_t)mx * my * mz;
  double *xl = (double *)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) xl[i] = (mask && mask[i]) ? 0.0 : u[i];
  memset(yp, 0, sizeof(double) * np);
  ptq2o_a21(mx, my, mz, coords, xl, yp);
  free(xl);
}

/* MatGetDiagonal_MFStokes_A11 (src/stokes_operators.c:1072-1113): Dirichlet slots := 1 */
void ptq2o_matgetdiag_a11(int mx, int my, int mz, const double *coords, const double *eta,
                          const uint8_t *mask, double *d) {
  int64_t n = nvel(mx, my, mz);
  memset(d, 0, sizeof(double) * n);
  ptq2o_diag_a11(mx, my, mz, coords, eta, d);
  if (mask)
    for (int64_t i = 0; i < n; i++)
      if (mask[i]) d[i] = 1.0;
}

/* MatMult_MFStokes_A (src/stokes_operators.c:893-979) */
void ptq2o_matmult_a(int mx, int my, int mz, const double *coords, const double *eta,
                     const uint8_t *mask, const double *u, const double *p, double *yu,
                     double *yp) {
  int64_t n = nvel(mx, my, mz), np = 4 * (int64_t)mx * my * mz;
  double *xl = (double *)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) xl[i] = (mask && mask[i]) ? 0.0 : u[i];
  memset(yu, 0, sizeof(double) * n);
  memset(yp, 0, sizeof(double) * np);
  ptq2o_stokes_a(mx, my, mz, coords, eta, xl, p, yu, yp);
  if (mask)
    for (int64_t i = 0; i < n; i++)
      if (mask[i]) yu[i] = u[i];
  free(xl);
}

/* ------------------------------------------------------------- residual */

/* FormFunction_Stokes (src/stokes_form_function.c:478-580) on one rank with the
   per-quadrature-point element loops of FormFunctionLocal_U (:99-242) and FormFunctionLocal_P
   (:365-463):
     u_loc = u with the Dirichlet values phi inserted (BCListInsertLocal, src/dmda_bcs.c:547-573);
     Fu = sum_e [ int dN . (2 eta eps(u_loc) - p I) - int N fu ]      (fu = QPntVolCoefStokes.Fu)
     Fp = sum_e [ -int N^p div u_loc - int N^p fp ]                  (fp = QPntVolCoefStokes.Fp)
     Fu_D = u_D - phi_D   (BCListResidualDirichlet, src/dmda_bcs.c:614-648, scale 1).
   fu: 81 per element ([e][q][d]), fp: 27 per element or NULL (0), phi: 3N or NULL (0). */
void ptq2o_residual(int mx, int my, int mz, const double *coords, const double *eta, const uint8_t *mask,
                    const double *phi, const double *u, const double *p, const double *fu, const double *fp,
                    double *Fu, double *Fp) {
  static double W[NQP], XI[NQP][3], NI[NQP][27], GNI[NQP][3][27];
  prepare_q2_3x3(W, XI, NI, GNI);
  const int64_t n = nvel(mx, my, mz), nel = (int64_t)mx * my * mz;
  double *ul = (double *)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) {
    ul[i] = (mask && mask[i]) ? (phi ? phi[i] : 0.0) : u[i];
    Fu[i] = 0.0;
  }
  for (int64_t e = 0; e < nel; e++) {
    double elx[3][27], elu[3][27], Fe[3][27], Be[3][27], Fpe[4] = {0, 0, 0, 0};
    gather_element(mx, my, (int)e, coords, elx);
    gather_element(mx, my, (int)e, ul, elu);
    memset(Fe, 0, sizeof Fe);
    memset(Be, 0, sizeof Be);
    for (int q = 0; q < NQP; q++) {
      double detJ, dN[3][27], Np[4];
      geometry_qp(GNI[q], elx, &detJ, dN);
      pressure_basis(XI[q], elx, Np);
      const double fac = W[q] * detJ, et = eta[27 * e + q];
      double pq = 0.0;
      for (int i = 0; i < 4; i++) pq += Np[i] * p[4 * e + i];
      pq *= fac;
      double exx = 0, eyy = 0, ezz = 0, exy = 0, exz = 0, eyz = 0;
      for (int k = 0; k < 27; k++) {
        exx += dN[0][k] * elu[0][k];
        eyy += dN[1][k] * elu[1][k];
        ezz += dN[2][k] * elu[2][k];
        exy += dN[1][k] * elu[0][k] + dN[0][k] * elu[1][k];
        exz += dN[2][k] * elu[0][k] + dN[0][k] * elu[2][k];
        eyz += dN[2][k] * elu[1][k] + dN[1][k] * elu[2][k];
      }
      const double d1 = 2.0 * et * fac, d2 = et * fac;
      const double sxx = d1 * exx, syy = d1 * eyy, szz = d1 * ezz, sxy = d2 * exy, sxz = d2 * exz, syz = d2 * eyz;
      for (int k = 0; k < 27; k++) {
        Fe[0][k] += dN[0][k] * (sxx - pq) + dN[1][k] * sxy + dN[2][k] * sxz;
        Fe[1][k] += dN[1][k] * (syy - pq) + dN[0][k] * sxy + dN[2][k] * syz;
        Fe[2][k] += dN[2][k] * (szz - pq) + dN[0][k] * sxz + dN[1][k] * syz;
        for (int d = 0; d < 3; d++) Be[d][k] += fac * NI[q][k] * fu[81 * e + 3 * q + d];
      }
      const double div = -(exx + eyy + ezz) * fac;
      const double f = fp ? fp[27 * e + q] : 0.0;
      for (int i = 0; i < 4; i++) Fpe[i] += Np[i] * div - fac * Np[i] * f;
    }
    for (int d = 0; d < 3; d++)
      for (int k = 0; k < 27; k++) Fe[d][k] -= Be[d][k];
    scatter_element(mx, my, (int)e, Fe, Fu);
    for (int i = 0; i < 4; i++) Fp[4 * e + i] = Fpe[i];
  }
  if (mask)
    for (int64_t i = 0; i < n; i++)
      if (mask[i]) Fu[i] = u[i] - (phi ? phi[i] : 0.0);
  free(ul);
}

/* The Be loop of FormFunctionLocal_U alone (src/stokes_form_function.c:219-224): y += sum_e
   sum_q w detJ N fu, per quadrature point (the checker of the device body-force load). */
void ptq2o_body_force(int mx, int my, int mz, const double *coords, const double *fu, double *y) {
  static double W[NQP], XI[NQP][3], NI[NQP][27], GNI[NQP][3][27];
  prepare_q2_3x3(W, XI, NI, GNI);
  const int64_t nel = (int64_t)mx * my * mz;
  for (int64_t e = 0; e < nel; e++) {
    double elx[3][27], Be[3][27];
    gather_element(mx, my, (int)e, coords, elx);
    memset(Be, 0, sizeof Be);
    for (int q = 0; q < NQP; q++) {
      double detJ, dN[3][27];
      geometry_qp(GNI[q], elx, &detJ, dN);
      const double fac = W[q] * detJ;
      for (int k = 0; k < 27; k++)
        for (int d = 0; d < 3; d++) Be[d][k] += fac * NI[q][k] * fu[81 * e + 3 * q + d];
    }
    scatter_element(mx, my, (int)e, Be, y);
  }
}

/* ---------------------------------------------------------------- markers */

/* pTatin_ConstructNi_Q2_3D (src/element_type_Q2.c:47-69): N[i+3j+9k] = l_i(xi) l_j(eta) l_k(zeta) */
static void q2_basis_point(const double xi[3], double N[27]) {
  double b[3][3];
  for (int d = 0; d < 3; d++) {
    const double x = xi[d];
    b[d][0] = 0.5 * x * (x - 1.0);
    b[d][1] = (1.0 + x) * (1.0 - x);
    b[d][2] = 0.5 * (1.0 + x) * x;
  }
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) N[i + 3 * j + 9 * k] = b[0][i] * b[1][j] * b[2][k];
}

/* pTatinConstructNI_Q1_3D (src/material_point_utils.c:503-520): vertex order x fastest */
static void q1_basis_point(const double xi[3], double N[8]) {
  for (int c = 0; c < 8; c++) {
    const double sx = (c & 1) ? 1.0 : -1.0, sy = (c & 2) ? 1.0 : -1.0, sz = (c & 4) ? 1.0 : -1.0;
    N[c] = 0.125 * (1.0 + sx * xi[0]) * (1.0 + sy * xi[1]) * (1.0 + sz * xi[2]);
  }
}

/* the Q2 element node of Q1 vertex c: 0, 2, 6, 8, 18, 20, 24, 26 */
static int q1_vertex_node(int c) { return 2 * (c & 1) + 6 * ((c >> 1) & 1) + 18 * ((c >> 2) & 1); }

/* SwarmMPntStd_CoordAssignment_LatticeLayout3d (src/material_point_std_utils.c:100-210) with a
   zero perturbation: nx*ny*nz points per element, element-major, x fastest, at
   xi_d = -1 + (2/n_d)(p_d + 1/2); wil = element; the physical coordinate is the Q2 interpolation
   of the element's nodes at xi (MaterialPointCoordinateSetUp, src/ptatin3d.c:520-562, which
   re-locates the points after the mesh geometry is applied). */
void ptq2o_marker_lattice(int mx, int my, int mz, const double *coords, int nx, int ny, int nz, double *xi,
                          int32_t *wil, double *xp) {
  const int64_t nel = (int64_t)mx * my * mz;
  const double dxi = 2.0 / nx, deta = 2.0 / ny, dzeta = 2.0 / nz;
  int64_t p = 0;
  for (int64_t e = 0; e < nel; e++) {
    double elx[3][27];
    gather_element(mx, my, (int)e, coords, elx);
    for (int pk = 0; pk < nz; pk++)
      for (int pj = 0; pj < ny; pj++)
        for (int pi = 0; pi < nx; pi++, p++) {
          double x[3] = {-1.0 + dxi * (pi + 0.5), -1.0 + deta * (pj + 0.5), -1.0 + dzeta * (pk + 0.5)};
          double N[27];
          q2_basis_point(x, N);
          for (int d = 0; d < 3; d++) {
            double s = 0.0;
            for (int k = 0; k < 27; k++) s += N[k] * elx[d][k];
            xp[3 * p + d] = s;
            xi[3 * p + d] = x[d];
          }
          wil[p] = (int32_t)e;
        }
  }
}

/* SwarmUpdateGaussPropertiesLocalL2Projection_Q1_MPntPStokes (src/material_point_utils.c:905-955)
   through _SwarmUpdateGaussPropertiesLocalL2ProjectionQ1_MPntPStokes (:663-830), one rank:
   1. every point adds w_i*eta_p, w_i*rho_p and w_i into the 27 nodes of its element, with w_i the
      Q1 basis at the point for the 8 vertex nodes and 1 for the other 19 nodes (:705-740);
   2. nodal values = sums / weight sums (VecPointwiseDivide, :760-761; a zero weight gives 0);
   3. each quadrature point (3-point Gauss rule, truncated constants of QuadratureCreateGauss_3pnt_3D,
      src/quadrature.c:136-167, q = i+3j+9k) takes the Q1 interpolation of the element's 8 vertex
      values (:766-782,816-824), clipped to [min, max] of the element's 27 nodal values (:804-829).
   rho_p / rho_qp may be NULL. */
void ptq2o_project_q1(int mx, int my, int mz, int64_t npoints, const double *xi, const int32_t *wil,
                      const double *eta_p, const double *rho_p, double *eta_qp, double *rho_qp) {
  const int NX = 2 * mx + 1, NY = 2 * my + 1, NZ = 2 * mz + 1;
  const int64_t N = (int64_t)NX * NY * NZ, nel = (int64_t)mx * my * mz;
  double *A1 = (double *)calloc((size_t)N, sizeof(double));
  double *A2 = (double *)calloc((size_t)N, sizeof(double));
  double *Bw = (double *)calloc((size_t)N, sizeof(double));
  for (int64_t p = 0; p < npoints; p++) {
    double w[27], n1[8];
    q1_basis_point(&xi[3 * p], n1);
    for (int i = 0; i < 27; i++) w[i] = 1.0;
    for (int c = 0; c < 8; c++) w[q1_vertex_node(c)] = n1[c];
    const int e = wil[p];
    const int ex = e % mx, ey = (e / mx) % my, ez = e / (mx * my);
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
          const int64_t n = nodeid(NX, NY, 2 * ex + i, 2 * ey + j, 2 * ez + k);
          const double wi = w[i + 3 * j + 9 * k];
          A1[n] += wi * eta_p[p];
          if (rho_p) A2[n] += wi * rho_p[p];
          Bw[n] += wi;
        }
  }
  for (int64_t n = 0; n < N; n++) {
    A1[n] = Bw[n] != 0.0 ? A1[n] / Bw[n] : 0.0;
    A2[n] = Bw[n] != 0.0 ? A2[n] / Bw[n] : 0.0;
  }
  const double s = 0.774596669241483, x1[3] = {-s, 0.0, s};
  double NQ[NQP][8];
  for (int q = 0; q < NQP; q++) {
    double x[3] = {x1[q % 3], x1[(q / 3) % 3], x1[q / 9]};
    q1_basis_point(x, NQ[q]);
  }
  for (int64_t e = 0; e < nel; e++) {
    const int ex = (int)(e % mx), ey = (int)((e / mx) % my), ez = (int)(e / ((int64_t)mx * my));
    double a1[27], a2[27];
    double lo1 = 1.0e32, hi1 = -1.0e32, lo2 = 1.0e32, hi2 = -1.0e32;
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
          const int64_t n = nodeid(NX, NY, 2 * ex + i, 2 * ey + j, 2 * ez + k);
          const int l = i + 3 * j + 9 * k;
          a1[l] = A1[n];
          a2[l] = A2[n];
          if (a1[l] < lo1) lo1 = a1[l];
          if (a1[l] > hi1) hi1 = a1[l];
          if (a2[l] < lo2) lo2 = a2[l];
          if (a2[l] > hi2) hi2 = a2[l];
        }
    for (int q = 0; q < NQP; q++) {
      double v1 = 0.0, v2 = 0.0;
      for (int c = 0; c < 8; c++) {
        v1 += NQ[q][c] * a1[q1_vertex_node(c)];
        v2 += NQ[q][c] * a2[q1_vertex_node(c)];
      }
      if (v1 < lo1) v1 = lo1;
      if (v1 > hi1) v1 = hi1;
      if (v2 < lo2) v2 = lo2;
      if (v2 > hi2) v2 = hi2;
      eta_qp[NQP * e + q] = v1;
      if (rho_qp) rho_qp[NQP * e + q] = v2;
    }
  }
  (void)NZ;
  free(A1);
  free(A2);
  free(Bw);
}
