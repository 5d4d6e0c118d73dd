// Occupied orbitals and density matrices on the FP64 matrix cores (density.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace gfn1 {

// C[l] = Cb[l] Z[l] (ld x kcap, column-major): Cb upper triangular (L^-T),
// Z element (m, j) at Z[l][m * zr + j * zc], of which rows >= n and columns
// >= kneed are zero.  One launch per bucket; act / count as for the other
// batched kernels.
hipError_t launch_orbitals(hipStream_t s, int count, const int* act, const int* nvec, const SccState* st, int ld,
                           int kcap, double* const* Cb, double* const* Z, int zr, int zc, double* const* C);

// M[l] = C diag(w) C^T on the n x n block, exactly symmetric, with w = occ
// (density) or eps * occ (weighted: energy-weighted density) over the first
// min(kneed, kcap) orbitals.  With rowS / rowH non-null, the Mulliken partial
// sums of S o M and H0 o M per 64-column block:
// rowS[blk * stride + ao0[inst] + mu], blk < density_row_blocks(n).
hipError_t launch_density(hipStream_t s, int count, const int* act, const int* nvec, const SccState* st, int ld,
                          int kcap, double* const* C, const double* eps_all, const double* occ_all,
                          const int64_t* vec0, bool weighted, double* const* M, double* const* S,
                          double* const* H0, const int* ao0, double* rowS, double* rowH, int stride);

// Number of 64-column partial blocks of the Mulliken sums of an n-AO matrix.
int density_row_blocks(int n);

}  // namespace gfn1
