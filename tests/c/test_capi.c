/* C consumer of libptq2 (include/ptq2.h, include/ptq2_solver.h): every entry point links,
   and the argument checks return codes instead of aborting.  Runs without a GPU; with a
   GPU present it also creates and destroys a context. */
#include <stdio.h>
#include <stdlib.h>

#include "ptq2.h"
#include "ptq2_solver.h"

#define EXPECT(cond) do { if (!(cond)) { fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, #cond); return 1; } } while (0)

int main(void) {
  ptq2_ctx *ctx = NULL;
  int rc = ptq2_create(&ctx, 0);
  EXPECT(rc == PTQ2_OK || rc == PTQ2_ERR_DEVICE);
  EXPECT(ptq2_create(NULL, 0) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_a11_apply(NULL, NULL, NULL, 0) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_a12_apply(NULL, NULL, NULL, 0) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_a21_apply(NULL, NULL, NULL, 0) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_stokes_apply(NULL, NULL, NULL, NULL, NULL, 0) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_a11_diag(NULL, NULL, 0) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_set_viscosity(NULL, NULL, 0, 0) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_setup_mesh(NULL, 1, 1, 1, NULL, NULL, NULL) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_setup_mesh_box(NULL, 1, 1, 1, 3, 3, 3, 0, 0, 0, NULL, NULL) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_set_stream(NULL, NULL) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_get_stream(NULL) == NULL);
  EXPECT(ptq2_sync(NULL) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_sizes(NULL, NULL, NULL, NULL) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_strerror(PTQ2_ERR_MESH) != NULL);
  EXPECT(ptq2_last_error(NULL) != NULL);
  ptq2_mg_opts o;
  ptq2_mg_default_opts(&o);
  EXPECT(o.nlevels >= 2 && o.smooth_its == 4 && o.fp32_smoother == 1);
  ptq2_mg *mg = NULL;
  EXPECT(ptq2_mg_create(&mg, NULL, &o) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_mg_vcycle(NULL, NULL, NULL) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_mg_solve(NULL, NULL, NULL, 1e-8, 10, 30, NULL, NULL, NULL) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_mg_levels(NULL, NULL, NULL) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_mg_eigs(NULL, NULL, NULL) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_mg_history(NULL, NULL, 0, NULL) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_mg_set_stream(NULL, NULL) == PTQ2_ERR_INVALID_ARG);
  EXPECT(ptq2_mg_destroy(NULL) == PTQ2_OK);
  if (rc == PTQ2_OK) {
    double c[81] = {0};
    /* a mesh that does not fit its node box */
    EXPECT(ptq2_setup_mesh_box(ctx, 2, 1, 1, 3, 3, 3, 0, 0, 0, c, NULL) == PTQ2_ERR_INVALID_ARG);
    EXPECT(ptq2_a11_apply(ctx, c, c, 0) == PTQ2_ERR_NOT_SETUP);
    EXPECT(ptq2_destroy(&ctx) == PTQ2_OK && ctx == NULL);
  }
  printf("capi ok (context %s)\n", rc == PTQ2_OK ? "created" : "unavailable: no GPU");
  return 0;
}
