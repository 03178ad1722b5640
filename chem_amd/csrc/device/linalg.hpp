// Dense FP64 Cholesky factorisation and triangular inverse on the GPU
// (replace rocsolver_dpotrf / rocsolver_dtrtri, see linalg.hip).
#pragma once

#include "dev.hpp"

namespace mchip {

// A (n x n, column-major, lda) = L L^T in place (lower triangle; the strict
// upper triangle is left as it was).  Returns 0, or the 1-based index of the
// first non-positive pivot (LAPACK/rocSOLVER convention).  Synchronises the
// context's stream.
int chol_lower(Ctx& c, double* A, int n, int lda);

// In-place inverse of the lower-triangular, non-unit L (n x n, column-major,
// lda); the strict upper triangle is left as it was.
void trtri_lower(Ctx& c, double* A, int n, int lda);

}  // namespace mchip
