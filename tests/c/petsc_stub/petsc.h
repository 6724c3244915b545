/* Minimal stand-in for the PETSc declarations the -a11_op hip adapter uses: enough for a
   syntax and type check of ptatin-gene_amd/adapter/stokes_operators_hip.c in the CPU test
   suite (tests/test_capi.py).  Not a PETSc implementation; nothing here is linked or run. */
#ifndef PTQ2_TEST_PETSC_STUB_H
#define PTQ2_TEST_PETSC_STUB_H
#include <stdint.h>
#include <stddef.h>
typedef int PetscErrorCode;
typedef int PetscInt;
typedef int PetscMPIInt;
typedef double PetscReal;
typedef double PetscScalar;
typedef int64_t PetscObjectState;
typedef enum { PETSC_FALSE = 0, PETSC_TRUE = 1 } PetscBool;
typedef int MPI_Comm;
typedef struct _p_DM *DM;
typedef struct _p_Vec *Vec;
typedef struct _p_PetscOptions *PetscOptions;
#define PETSC_COMM_SELF ((MPI_Comm)1)
#define PETSC_COMM_WORLD ((MPI_Comm)2)
#define PETSC_ERR_LIB 76
#define PETSC_ERR_ORDER 73
#define PetscFunctionBegin do { } while (0)
#define PetscFunctionReturn(x) return (x)
#define CHKERRQ(e) do { if (e) return (e); } while (0)
PetscErrorCode PetscError(MPI_Comm, int, const char *, const char *, int, int, const char *, ...);
#define SETERRQ(c, n, s) return PetscError((c), __LINE__, __func__, __FILE__, (n), 1, (s))
#define SETERRQ2(c, n, s, a1, a2) return PetscError((c), __LINE__, __func__, __FILE__, (n), 1, (s), (a1), (a2))
PetscErrorCode PetscMallocA(int, PetscBool, int, const char *, const char *, size_t, void *, ...);
PetscErrorCode PetscFreeA(int, int, const char *, const char *, void *, ...);
#define PetscNew(b) PetscMallocA(1, PETSC_TRUE, __LINE__, __func__, __FILE__, sizeof(**(b)), (b))
#define PetscMalloc1(m, r) PetscMallocA(1, PETSC_FALSE, __LINE__, __func__, __FILE__, (size_t)(m) * sizeof(**(r)), (r))
#define PetscFree(a) (PetscFreeA(1, __LINE__, __func__, __FILE__, &(a)) || ((a) = NULL, 0))
int MPI_Comm_rank(MPI_Comm, PetscMPIInt *);
PetscErrorCode PetscOptionsGetInt(PetscOptions, const char[], const char[], PetscInt *, PetscBool *);
PetscErrorCode DMGetCoordinatesLocal(DM, Vec *);
PetscErrorCode VecGetArrayRead(Vec, const PetscScalar **);
PetscErrorCode VecRestoreArrayRead(Vec, const PetscScalar **);
#endif
