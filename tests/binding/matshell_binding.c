/*
 * The reference-side binding of the MatShell callbacks, compiled as C11
 * against include/laph.h: MatMult_Laplacian2D and MatGetDiagonal_Laplacian2D
 * with the signatures of modules/shell_matop.h:25-26 and the bodies a
 * maintainer would put in shell_matop.cpp (the repository's integration
 * notes show the same bodies), over minimal stand-ins for the PETSc handles
 * and calls they use (Mat carries the MatShell context, Vec a complex double
 * array).  Exported for the tests: binding_mat_mult / binding_get_diagonal
 * wrap a laph_ctx* and plain arrays in those handles and call the callbacks
 * as EPSSolve's MatMult would.
 */
#include <complex.h>
#include <stddef.h>

#include "laph.h"

/* --- PETSc stand-ins: just what the two callbacks touch ------------------ */
typedef int PetscErrorCode;
typedef double _Complex PetscScalar;
typedef struct { void* ctx; } MatObj;
typedef MatObj* Mat;
typedef struct { PetscScalar* a; } VecObj;
typedef VecObj* Vec;
#define PETSC_SUCCESS 0
#define PETSC_ERR_LIB 76
#define PETSC_COMM_SELF 0
#define PetscFunctionBeginUser do { } while (0)
#define PetscFunctionReturn(e) return (e)
#define PetscCall(e)                      \
  do {                                    \
    PetscErrorCode ierr_ = (e);           \
    if (ierr_ != PETSC_SUCCESS) return ierr_; \
  } while (0)
#define PetscCheck(cond, comm, code, ...) \
  do {                                    \
    if (!(cond)) return (code);           \
  } while (0)

static PetscErrorCode MatShellGetContext(Mat A, void* ctx) {
  *(void**)ctx = A->ctx;
  return PETSC_SUCCESS;
}
static PetscErrorCode VecGetArrayRead(Vec v, const PetscScalar** a) {
  *a = v->a;
  return PETSC_SUCCESS;
}
static PetscErrorCode VecRestoreArrayRead(Vec v, const PetscScalar** a) {
  (void)v;
  *a = NULL;
  return PETSC_SUCCESS;
}
static PetscErrorCode VecGetArray(Vec v, PetscScalar** a) {
  *a = v->a;
  return PETSC_SUCCESS;
}
static PetscErrorCode VecRestoreArray(Vec v, PetscScalar** a) {
  (void)v;
  *a = NULL;
  return PETSC_SUCCESS;
}

/* --- the callbacks (modules/shell_matop.cpp:226-255 replaced) ------------ */
PetscErrorCode MatMult_Laplacian2D(Mat A, Vec x, Vec y) {
  void* ctx;
  const PetscScalar* px;
  PetscScalar* py;
  PetscFunctionBeginUser;
  PetscCall(MatShellGetContext(A, &ctx));
  PetscCall(VecGetArrayRead(x, &px));
  PetscCall(VecGetArray(y, &py));
  int rc = laph_shell_mult(ctx, (const double*)px, (double*)py); /* y = T_DEG(B) x on the GPU */
  PetscCall(VecRestoreArrayRead(x, &px));
  PetscCall(VecRestoreArray(y, &py));
  PetscCheck(rc == 0, PETSC_COMM_SELF, PETSC_ERR_LIB, "laph_shell_mult: %s", laph_last_error());
  PetscFunctionReturn(PETSC_SUCCESS);
}

PetscErrorCode MatGetDiagonal_Laplacian2D(Mat A, Vec diag) {
  void* ctx;
  PetscScalar* pd;
  PetscFunctionBeginUser;
  PetscCall(MatShellGetContext(A, &ctx));
  PetscCall(VecGetArray(diag, &pd));
  int rc = laph_shell_get_diagonal(ctx, (double*)pd); /* -6, as shell_matop.cpp:253 */
  PetscCall(VecRestoreArray(diag, &pd));
  PetscCheck(rc == 0, PETSC_COMM_SELF, PETSC_ERR_LIB, "%s", laph_last_error());
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* --- entry points for the tests ------------------------------------------ */
int binding_mat_mult(void* laph_context, const double* x, double* y) {
  MatObj A = {laph_context};
  VecObj vx = {(PetscScalar*)x}, vy = {(PetscScalar*)y};
  return MatMult_Laplacian2D(&A, &vx, &vy);
}

int binding_get_diagonal(void* laph_context, double* d) {
  MatObj A = {laph_context};
  VecObj vd = {(PetscScalar*)d};
  return MatGetDiagonal_Laplacian2D(&A, &vd);
}
