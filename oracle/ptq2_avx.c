/*
 * ptq2_avx.c -- CPU BASELINE, test and measurement infrastructure only (loaded by tests/ and by
 * bench.py's cpu_baseline leg; the product library never loads it).
 *
 * Restatement of pTatin3D's default CPU A11 backend, MFStokesWrapper_A11_AVX
 * (src/stokes_operators_avx.c:206-312): elements in batches of NEV = 4 with one element per
 * __m256d lane, the 27 nodes x 3 components gathered per batch, nine 3-stage sum-factorized
 * 3x3x3 contractions (TensorContractNEV_AVX, :52-113), the cofactor Jacobian inverse
 * (JacobianInvertNEV_AVX, :115-146) and the quadrature action (QuadratureAction_A11_AVX,
 * :148-204), then the scatter-add.  The reference serialises the scatter with `omp critical`
 * (:289-291) and scales with MPI ranks instead; this restatement keeps the scatter race-free by
 * colouring (elements of one (ex,ey,ez) parity class share no node, a batch is 4 elements of one
 * class), so its OpenMP scaling is that of the arithmetic.  The quadrature viscosity is read
 * from the SoA array eta[27e+q] rather than from the 48-byte QPntVolCoefStokes records.
 * Built with -O3 -mavx2 -mfma (the reference's __m256d path on any x86-64 with AVX2/FMA).
 */
#include <immintrin.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>

#define NEV 4
#define NQP 27
#define AL __attribute__((aligned(32)))

typedef double blk[3][NQP][NEV]; /* [component][node or point][element of the batch] */

/* apply the 3x3 matrix M along axis `ax` of each component's 3x3x3 block (index i + 3j + 9k),
   accumulating when acc != 0:  out[..a..] (+)= sum_n M[a][n] in[..n..] */
static inline void axis_apply(const double M[3][3], int ax, const double in[3][NQP][NEV], double out[3][NQP][NEV],
                              int acc) {
  const int st = ax == 0 ? 1 : (ax == 1 ? 3 : 9);
  for (int c = 0; c < 3; c++)
    for (int p = 0; p < 9; p++) {
      /* the two other axes' indices of line p */
      int base;
      if (ax == 0) base = 3 * p;
      else if (ax == 1) base = (p % 3) + 9 * (p / 3);
      else base = p;
      const __m256d v0 = _mm256_load_pd(in[c][base]), v1 = _mm256_load_pd(in[c][base + st]),
                    v2 = _mm256_load_pd(in[c][base + 2 * st]);
      for (int a = 0; a < 3; a++) {
        __m256d s = acc ? _mm256_load_pd(out[c][base + a * st]) : _mm256_setzero_pd();
        s = _mm256_fmadd_pd(_mm256_set1_pd(M[a][0]), v0, s);
        s = _mm256_fmadd_pd(_mm256_set1_pd(M[a][1]), v1, s);
        s = _mm256_fmadd_pd(_mm256_set1_pd(M[a][2]), v2, s);
        _mm256_store_pd(out[c][base + a * st], s);
      }
    }
}

/* y (+)= (Mz (x) My (x) Mx) x over the three components: x-, y-, then z-axis stage */
static inline void contract3(const double Mx[3][3], const double My[3][3], const double Mz[3][3], const blk x, blk y,
                             int acc) {
  blk t1 AL, t2 AL;
  axis_apply(Mx, 0, x, t1, 0);
  axis_apply(My, 1, t1, t2, 0);
  axis_apply(Mz, 2, t2, y, acc);
}

static void transpose33(const double A[3][3], double T[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) T[i][j] = A[j][i];
}

void ptq2a_a11_avx(int mx, int my, int mz, const double *coords, const double *eta, const double *u, double *y,
                   int nthreads) {
  double B[3][3], D[3][3], Bt[3][3], Dt[3][3], w[NQP];
  const double g[3] = {-sqrt(0.6), 0.0, sqrt(0.6)}, w1[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  for (int q = 0; q < 3; q++) {
    B[q][0] = 0.5 * (g[q] * g[q] - g[q]);
    B[q][1] = 1.0 - g[q] * g[q];
    B[q][2] = 0.5 * (g[q] * g[q] + g[q]);
    D[q][0] = g[q] - 0.5;
    D[q][1] = -2.0 * g[q];
    D[q][2] = g[q] + 0.5;
  }
  transpose33(B, Bt);
  transpose33(D, Dt);
  for (int c = 0; c < 3; c++)
    for (int b = 0; b < 3; b++)
      for (int a = 0; a < 3; a++) w[a + 3 * b + 9 * c] = w1[a] * w1[b] * w1[c];
  const int64_t NX = 2 * mx + 1, NY = 2 * my + 1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  for (int col = 0; col < 8; col++) {
    const int cx = col & 1, cy = (col >> 1) & 1, cz = col >> 2;
    const int nxc = (mx - cx + 1) / 2, nyc = (my - cy + 1) / 2, nzc = (mz - cz + 1) / 2;
    const int nbx = (nxc + NEV - 1) / NEV; /* batches along x within the colour */
    const int64_t nbatch = (int64_t)nbx * nyc * nzc;
#pragma omp parallel for schedule(static)
    for (int64_t bt = 0; bt < nbatch; bt++) {
      const int bx = (int)(bt % nbx), ey = cy + 2 * (int)((bt / nbx) % nyc), ez = cz + 2 * (int)(bt / ((int64_t)nbx * nyc));
      int ex[NEV], ne = 0;
      for (int v = 0; v < NEV; v++) {
        int k = NEV * bx + v;
        if (k < nxc) ne = v + 1;
        ex[v] = cx + 2 * (k < nxc ? k : nxc - 1); /* pad the batch with its last element */
      }
      blk elx AL, elu AL, elv AL;
      double dx[3][3][NQP][NEV] AL, du[3][3][NQP][NEV] AL, dv[3][3][NQP][NEV] AL, det[NQP][NEV] AL;
      int64_t nid[NQP][NEV];
      for (int n = 0; n < NQP; n++)
        for (int v = 0; v < NEV; v++) {
          const int i = n % 3, j = (n / 3) % 3, k = n / 9;
          nid[n][v] = (2 * ex[v] + i) + NX * ((int64_t)(2 * ey + j) + NY * (int64_t)(2 * ez + k));
          for (int c = 0; c < 3; c++) {
            elx[c][n][v] = coords[3 * nid[n][v] + c];
            elu[c][n][v] = u[3 * nid[n][v] + c];
          }
        }
      /* reference gradients of the coordinates and of u: d/dxi, d/deta, d/dzeta */
      contract3(D, B, B, elx, dx[0], 0);
      contract3(B, D, B, elx, dx[1], 0);
      contract3(B, B, D, elx, dx[2], 0);
      contract3(D, B, B, elu, du[0], 0);
      contract3(B, D, B, elu, du[1], 0);
      contract3(B, B, D, elu, du[2], 0);
      /* cofactor inverse of J[d][l] = dx_l/dxi_d per point and element */
      for (int q = 0; q < NQP; q++)
        for (int v = 0; v < NEV; v++) {
          double a[3][3];
          for (int r = 0; r < 3; r++)
            for (int s = 0; s < 3; s++) a[r][s] = dx[r][s][q][v];
          const double c00 = a[1][1] * a[2][2] - a[2][1] * a[1][2];
          const double c10 = -(a[1][0] * a[2][2] - a[2][0] * a[1][2]);
          const double c20 = a[1][0] * a[2][1] - a[2][0] * a[1][1];
          const double dt = a[0][0] * c00 + a[0][1] * c10 + a[0][2] * c20, id = 1.0 / dt;
          dx[0][0][q][v] = id * c00;
          dx[0][1][q][v] = -id * (a[0][1] * a[2][2] - a[2][1] * a[0][2]);
          dx[0][2][q][v] = id * (a[0][1] * a[1][2] - a[1][1] * a[0][2]);
          dx[1][0][q][v] = id * c10;
          dx[1][1][q][v] = id * (a[0][0] * a[2][2] - a[2][0] * a[0][2]);
          dx[1][2][q][v] = -id * (a[0][0] * a[1][2] - a[1][0] * a[0][2]);
          dx[2][0][q][v] = id * c20;
          dx[2][1][q][v] = -id * (a[0][0] * a[2][1] - a[2][0] * a[0][1]);
          dx[2][2][q][v] = id * (a[0][0] * a[1][1] - a[1][0] * a[0][1]);
          det[q][v] = dt;
        }
      /* quadrature action: grad_x u = J^-1 grad_xi u, 2 eta sym(grad u), times w det, mapped back */
      for (int q = 0; q < NQP; q++) {
        __m256d gx[3][3];
        for (int k = 0; k < 3; k++)
          for (int l = 0; l < 3; l++) {
            __m256d s = _mm256_mul_pd(_mm256_load_pd(du[0][l][q]), _mm256_load_pd(dx[k][0][q]));
            s = _mm256_fmadd_pd(_mm256_load_pd(du[1][l][q]), _mm256_load_pd(dx[k][1][q]), s);
            gx[k][l] = _mm256_fmadd_pd(_mm256_load_pd(du[2][l][q]), _mm256_load_pd(dx[k][2][q]), s);
          }
        double e2[NEV] AL;
        for (int v = 0; v < NEV; v++) {
          const int64_t e = ex[v] + (int64_t)mx * (ey + (int64_t)my * ez);
          e2[v] = 2.0 * eta[27 * e + q];
        }
        const __m256d te = _mm256_load_pd(e2), h = _mm256_set1_pd(0.5);
        __m256d S[3][3];
        for (int k = 0; k < 3; k++) S[k][k] = _mm256_mul_pd(te, gx[k][k]);
        S[0][1] = S[1][0] = _mm256_mul_pd(te, _mm256_mul_pd(h, _mm256_add_pd(gx[0][1], gx[1][0])));
        S[0][2] = S[2][0] = _mm256_mul_pd(te, _mm256_mul_pd(h, _mm256_add_pd(gx[0][2], gx[2][0])));
        S[1][2] = S[2][1] = _mm256_mul_pd(te, _mm256_mul_pd(h, _mm256_add_pd(gx[1][2], gx[2][1])));
        const __m256d wd = _mm256_mul_pd(_mm256_set1_pd(w[q]), _mm256_load_pd(det[q]));
        for (int l = 0; l < 3; l++)
          for (int k = 0; k < 3; k++) {
            __m256d s = _mm256_mul_pd(S[0][l], _mm256_load_pd(dx[0][k][q]));
            s = _mm256_fmadd_pd(S[1][l], _mm256_load_pd(dx[1][k][q]), s);
            s = _mm256_fmadd_pd(S[2][l], _mm256_load_pd(dx[2][k][q]), s);
            _mm256_store_pd(dv[k][l][q], _mm256_mul_pd(wd, s));
          }
      }
      /* transposed contractions into the element output */
      contract3(Dt, Bt, Bt, dv[0], elv, 0);
      contract3(Bt, Dt, Bt, dv[1], elv, 1);
      contract3(Bt, Bt, Dt, dv[2], elv, 1);
      for (int v = 0; v < ne; v++)
        for (int n = 0; n < NQP; n++)
          for (int c = 0; c < 3; c++) y[3 * nid[n][v] + c] += elv[c][n][v];
    }
  }
}

/* the reference's flop count per element (src/stokes_operators_avx.c:305-307) */
double ptq2a_flops_per_element(void) { return 9.0 * 3 * NQP * 18 + NQP * (14 + 1 + 27) + NQP * (5 * 9 + 6 + 6 + 6 * 9); }
