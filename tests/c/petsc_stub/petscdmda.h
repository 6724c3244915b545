/* stand-in (see petsc.h in this folder) */
#include "petsc.h"
PetscErrorCode DMDAGetGhostCorners(DM, PetscInt *, PetscInt *, PetscInt *, PetscInt *, PetscInt *, PetscInt *);
