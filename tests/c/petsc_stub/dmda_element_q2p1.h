/* stand-in: element corners and connectivity of the Q2 DMDA */
#include "petsc.h"
PetscErrorCode DMDAGetCornersElementQ2(DM, PetscInt *, PetscInt *, PetscInt *, PetscInt *, PetscInt *, PetscInt *);
PetscErrorCode DMDAGetElements_pTatinQ2P1(DM, PetscInt *, PetscInt *, const PetscInt **);
