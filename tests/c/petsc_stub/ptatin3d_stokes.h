/* stand-in: the quadrature-point coefficients (only .eta is read) and their accessors */
#include "petsc.h"
typedef struct {
  double eta, rho, Fu[3], Fp;
} QPntVolCoefStokes;
typedef struct _p_Quadrature *Quadrature;
PetscErrorCode VolumeQuadratureGetAllCellData_Stokes(Quadrature, QPntVolCoefStokes **);
PetscErrorCode VolumeQuadratureGetCellData_Stokes(Quadrature, QPntVolCoefStokes *, PetscInt, QPntVolCoefStokes **);
