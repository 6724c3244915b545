/* stand-in: the A11 matrix-free context fields the SpMVOp hooks touch */
#include "petsc.h"
typedef struct _p_MatA11MF *MatA11MF;
struct _p_MatA11MF {
  PetscObjectState state;
  void *ctx;
};
