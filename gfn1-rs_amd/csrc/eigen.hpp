// Batched eigensolver for the SCC Fock matrices (eigen.hip).
//
// A batch of symmetric n x n matrices (n <= 512, column-major, leading
// dimension ld, one per active instance) is reduced to tridiagonal form, all
// eigenvalues are found by bisection, Fermi occupations are formed, and the
// eigenvectors of the occupied orbitals are computed by inverse iteration and
// transformed back.  The eigenvectors are returned transposed: Zt[r * kcap + j]
// is row r of eigenvector j (so the batched GEMM C = Cb Zt^T needs no copy).
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace gfn1 {

constexpr int kEigMaxDim = 512;     // largest matrix the kernels accept
constexpr int kEigPanel = 8;        // Householder panel width
constexpr int kBacktransformDefaultCols = 32;
constexpr int kEigMaxCluster = 32;  // largest tight eigenvalue cluster resolved by Rayleigh-Ritz
constexpr int kEigMargin = 16;      // eigenvalues formed above the occupied ones (partial spectrum)

// Householder reduction of the lower triangle of A[l] (n = nvec[act[l]]).  On
// exit d, e hold the tridiagonal matrix, A[l] column k rows k+1.. holds the
// reflector v_k (v_k[k+1] = 1 stored explicitly) and tau[k] its factor.
hipError_t launch_tridiag(hipStream_t s, int count, const int* act, const int* nvec, int ld, double* const* A,
                          double* d, double* e, double* tau);

// Eigenvalues of the tridiagonal matrices, occupations (writes eps, occ, mu, ts,
// kneed into the instance state) and the first min(kneed, kcap) eigenvectors of
// T into Zt[l] (transposed, rows n..ld-1 and columns kneed..kcap-1 zero).
// full_spectrum: every orbital energy is formed; otherwise the lowest
// nocc + kEigMargin (more if the Fermi window needs them), the others being
// reported as an upper bound with the exact occupation and entropy of an orbital
// 80 kT or more above the Fermi level.
// lu: 5 * ld * kcap doubles per active instance.
hipError_t launch_tridiag_eig(hipStream_t s, int count, const int* act, const int* nvec, int ld, const double* d,
                              const double* e, double* eps_all, const int64_t* vec0, double* occ_all, SccState* st,
                              double* const* Zt, double* lu, int kcap, bool full_spectrum);

// Zt[l] <- (Q Z)^T with Q the product of the reflectors left by launch_tridiag.
// work: 16 * ld doubles per active instance (the blocks' T factors); the lu
// workspace of launch_tridiag_eig is free at this point and large enough.
// cols: eigenvectors per work-group, 16, 32 or (ld <= 256) 64; every width
// gives the same result bits.
hipError_t launch_backtransform(hipStream_t s, int count, const int* act, const int* nvec, int ld,
                                double* const* A, const double* tau, double* const* Zt, const SccState* st,
                                int kcap, double* work, int cols = kBacktransformDefaultCols);
bool backtransform_width_ok(int ld, int cols);

}  // namespace gfn1
