/*
 * stokes_operators_hip.c -- the `-a11_op hip` backend of pTatin3D's A11 MatShell.
 *
 * Drop-in for the SpMVOp seam of struct _p_MatA11MF (include/stokes_operators.h:48-64):
 *   SpMVOp_SetUp   = MFA11SetUp_HIP      (called from MatA11MFSetup, src/stokes_operators.c:271)
 *   SpMVOp_MatMult = MFStokesWrapper_A11_HIP (called from MatMult_MFStokes_A11, :1199)
 *   SpMVOp_Destroy = MFA11Destroy_HIP    (called from MatA11MFDestroy, :329)
 * and registered next to the other backends in MatA11MFCreate (src/stokes_operators.c:123-142):
 *   PetscFunctionListAdd(&SpMVOp_MatMult_List,"hip",MFStokesWrapper_A11_HIP);
 *   PetscFunctionListAdd(&SpMVOp_SetUp_List,  "hip",MFA11SetUp_HIP);
 *   PetscFunctionListAdd(&SpMVOp_Destroy_List,"hip",MFA11Destroy_HIP);
 *
 * The reference's MatShell keeps doing the global-to-local scatter, the Dirichlet column
 * zeroing, the surface-BC hook, the local-to-global ADD and the Dirichlet row clobber; this
 * backend receives the ghosted local velocity `ufield` and adds the element contributions
 * into the zeroed local `Yu` -- the raw contract of the CUDA backend
 * (src/stokes_operators_cuda.cu:457-532) -- through libptq2's C ABI (include/ptq2.h).
 * Coordinates and the quadrature-point viscosity are uploaded when mf->state changes and
 * stay resident on the GPU; only u and Yu cross PCIe per call.  libptq2 returns error codes;
 * they are turned into PETSc errors here (no exit()).
 *
 * Build: part of a PETSc build of pTatin3D only (needs petsc.h and the pTatin3D headers);
 * link with -lptq2.
 */
#include <petsc.h>
#include <petscdmda.h>

#include "ptatin3d.h"
#include "ptatin3d_stokes.h"
#include "dmda_element_q2p1.h"
#include "stokes_operators.h"

#include "ptq2.h"

typedef struct {
  ptq2_ctx         *dev;
  PetscObjectState state;
  PetscBool        uploaded;
  PetscReal        *eta;   /* host staging of the 27 eta values per element */
  PetscInt         nel;
} MFA11HIP;

#define PTQ2CHK(rc, dev) do { int rc_ = (rc); \
  if (rc_) SETERRQ2(PETSC_COMM_SELF,PETSC_ERR_LIB,"libptq2: %s (%s)",ptq2_strerror(rc_),ptq2_last_error(dev)); } while (0)

PetscErrorCode MFA11SetUp_HIP(MatA11MF mf)
{
  PetscErrorCode ierr;
  MFA11HIP       *ctx;
  PetscMPIInt    rank;
  int            ndev = 1, device;
  PetscInt       opt = -1;
  PetscBool      flg;

  PetscFunctionBegin;
  if (mf->ctx) PetscFunctionReturn(0);
  ierr = PetscNew(&ctx);CHKERRQ(ierr);
  ierr = MPI_Comm_rank(PETSC_COMM_WORLD,&rank);CHKERRQ(ierr);
  ierr = PetscOptionsGetInt(NULL,NULL,"-a11_op_hip_device",&opt,&flg);CHKERRQ(ierr);
  device = flg ? (int)opt : (int)(rank % ndev); /* one rank per GPU: set HIP_VISIBLE_DEVICES per rank */
  PTQ2CHK(ptq2_create(&ctx->dev,device),NULL);
  mf->ctx = (void*)ctx;
  PetscFunctionReturn(0);
}

PetscErrorCode MFA11Destroy_HIP(MatA11MF mf)
{
  MFA11HIP       *ctx = (MFA11HIP*)mf->ctx;
  PetscErrorCode ierr;

  PetscFunctionBegin;
  if (!ctx) PetscFunctionReturn(0);
  ptq2_destroy(&ctx->dev);
  ierr = PetscFree(ctx->eta);CHKERRQ(ierr);
  ierr = PetscFree(ctx);CHKERRQ(ierr);
  mf->ctx = NULL;
  PetscFunctionReturn(0);
}

/* upload the rank-local ghosted node box, its coordinates and eta (on state change) */
static PetscErrorCode MFA11HIP_Upload(MFA11HIP *ctx,Quadrature volQ,DM dau,PetscObjectState state)
{
  PetscErrorCode    ierr;
  PetscInt          esi,esj,esk,mx,my,mz,gsi,gsj,gsk,X,Y,Z,e,q,nel,nen;
  const PetscInt    *elnidx;
  Vec               gcoords;
  const PetscScalar *LA_gcoords;
  QPntVolCoefStokes *all_gausspoints,*cell;

  PetscFunctionBegin;
  ierr = DMDAGetCornersElementQ2(dau,&esi,&esj,&esk,&mx,&my,&mz);CHKERRQ(ierr);
  ierr = DMDAGetGhostCorners(dau,&gsi,&gsj,&gsk,&X,&Y,&Z);CHKERRQ(ierr);
  ierr = DMDAGetElements_pTatinQ2P1(dau,&nel,&nen,&elnidx);CHKERRQ(ierr);
  ierr = DMGetCoordinatesLocal(dau,&gcoords);CHKERRQ(ierr);
  ierr = VecGetArrayRead(gcoords,&LA_gcoords);CHKERRQ(ierr);
  PTQ2CHK(ptq2_setup_mesh_box(ctx->dev,(int)mx,(int)my,(int)mz,(int)X,(int)Y,(int)Z,
                              (int)(esi-gsi),(int)(esj-gsj),(int)(esk-gsk),LA_gcoords,NULL),ctx->dev);
  ierr = VecRestoreArrayRead(gcoords,&LA_gcoords);CHKERRQ(ierr);

  if (ctx->nel != nel) {
    ierr = PetscFree(ctx->eta);CHKERRQ(ierr);
    ierr = PetscMalloc1(27*nel,&ctx->eta);CHKERRQ(ierr);
    ctx->nel = nel;
  }
  ierr = VolumeQuadratureGetAllCellData_Stokes(volQ,&all_gausspoints);CHKERRQ(ierr);
  for (e=0; e<nel; e++) {
    ierr = VolumeQuadratureGetCellData_Stokes(volQ,all_gausspoints,e,&cell);CHKERRQ(ierr);
    for (q=0; q<27; q++) ctx->eta[27*e+q] = cell[q].eta;
  }
  PTQ2CHK(ptq2_set_viscosity(ctx->dev,ctx->eta,(uint64_t)state,0),ctx->dev);
  ctx->state    = state;
  ctx->uploaded = PETSC_TRUE;
  PetscFunctionReturn(0);
}

PetscErrorCode MFStokesWrapper_A11_HIP(MatA11MF mf,Quadrature volQ,DM dau,PetscScalar ufield[],PetscScalar Yu[])
{
  MFA11HIP       *ctx = (MFA11HIP*)mf->ctx;
  PetscErrorCode ierr;

  PetscFunctionBegin;
  if (!ctx) SETERRQ(PETSC_COMM_SELF,PETSC_ERR_ORDER,"MFA11SetUp_HIP was not called");
  if (!ctx->uploaded || ctx->state != mf->state) {
    ierr = MFA11HIP_Upload(ctx,volQ,dau,mf->state);CHKERRQ(ierr);
  }
  /* Yu += sum of element contributions, host arrays staged through the device */
  PTQ2CHK(ptq2_a11_apply(ctx->dev,ufield,Yu,PTQ2_ADD),ctx->dev);
  PetscFunctionReturn(0);
}
