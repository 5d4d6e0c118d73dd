// Householder tridiagonalisation of one large symmetric matrix over the whole
// device (tridiag_large.hip), for the matrices above the batched eigensolver's
// 512 rows (single large systems).
#pragma once

#include <hip/hip_runtime.h>

namespace gfn1 {

constexpr int kLargeTridiagPanel = 32;       // reflectors per trailing-matrix update (LAPACK's nb)
constexpr int kLargeTridiagMaxDim = 16384;   // rows one column work-group covers (16 per thread)

// Scratch of launch_tridiag_large for matrices of up to n rows (device memory).
struct LargeTridiagWork {
    int n = 0;
    double* V = nullptr;     // n x kLargeTridiagPanel reflectors of the current panel
    double* W = nullptr;     // n x kLargeTridiagPanel
    double* part = nullptr;  // ceil(n / 64) x n partial sums of the matrix-vector product
    double* dots = nullptr;  // per 256-row chunk: V^T u, W^T u and |u|^2
    double* pv = nullptr;    // per 256-row chunk: p . v, two columns' worth
    double* next = nullptr;  // the next column after the panel's corrections but the last (n)
    double* scal = nullptr;  // the current column's scalars and the panel's a2 values
    hipError_t allocate(int rows);
    void release();
};

// LAPACK dsytrd (uplo = 'L') of the n x n matrix A (column-major, leading
// dimension ld, lower triangle read): on exit d[0..n-1] and e[0..n-2] hold the
// tridiagonal matrix, tau[0..n-2] the reflector factors, and column k of A
// below row k+1 the reflector v_k (v_k[k+1] = 1 implicit, A(k+1, k) = e[k]), the
// layout rocsolver_dormtr reads.  Panels of kLargeTridiagPanel columns (dlatrd)
// with the rank-2 kLargeTridiagPanel trailing update on the FP64 matrix
// cores, all on stream s.
hipError_t launch_tridiag_large(hipStream_t s, int n, int ld, double* A, double* d, double* e, double* tau,
                                LargeTridiagWork& w);

}  // namespace gfn1
