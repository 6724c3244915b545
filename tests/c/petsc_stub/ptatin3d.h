/* stand-in for the pTatin3D declarations the adapter uses (see petsc.h in this folder) */
#include "petsc.h"
