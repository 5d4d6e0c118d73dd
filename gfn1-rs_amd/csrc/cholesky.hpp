// Batched Cholesky factorisation and orthonormal start basis (cholesky.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace gfn1 {

constexpr int kCholMaxDim = 512;  // largest ld the kernels accept (a multiple of 32)

// For every listed matrix l (ld x ld, column-major, n = nvec[act[l]] real rows,
// padding block the identity): L[l] = chol(S[l]) (lower, strict upper zero),
// Cb[l] = L[l]^-T (upper, zero below the diagonal), fail[l] = 1 when S[l] is
// not positive definite or a pivot of its first n rows is <= eig_tol, else 0.
hipError_t launch_cholesky_cb(hipStream_t s, int count, const int* act, const int* nvec, int ld, double* const* S,
                              double* const* L, double* const* Cb, double eig_tol, int* fail);

}  // namespace gfn1
