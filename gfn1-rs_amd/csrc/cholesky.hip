// Batched Cholesky factorisation S = L L^T and orthonormal start basis
// Cb = L^-T of the overlap matrices (gfx950), one work-group per matrix.
//
// The reference builds the orthonormalising transformation from S in its
// generalized eigensolver (linalg.rs:153-191, 315-325); the engine keeps L for
// the reduced congruence of every SCC iteration and Cb for the orbitals.
//
//   k_cholesky    right-looking blocked Cholesky, panels of 16 columns: the
//                 panel is factored in LDS (one barrier per column), the
//                 trailing lower triangle is updated in 16 x 16 tiles on the
//                 FP64 matrix cores
//   k_triinv_cb   X = L^-1 one 16-row block at a time,
//                 X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj, both products on the
//                 matrix cores; X is written transposed, i.e. as Cb
#include <hip/hip_runtime.h>

#include <cmath>

#include "cholesky.hpp"
#include "device.hpp"

namespace gfn1 {
namespace {

constexpr int kCT = 512;      // threads per work-group
constexpr int kCW = kCT / 64;  // wavefronts
constexpr int kCB = 16;        // panel / block width
constexpr int kPS = kCB + 1;   // LDS row stride of the staged panel

typedef double dvec4 __attribute__((ext_vector_type(4)));

// D = A B + C on a 16 x 16 x 4 tile: A[i][k] in lane i + 16k, B[k][j] in lane
// j + 16k, element e of D in lane l is D[4e + l/16][l%16].
__device__ __forceinline__ dvec4 mfma4(double a, double b, dvec4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

}  // namespace

// L = chol(S) (lower, strict upper zero) over the whole ld x ld matrix; the
// padding block of S is the identity.  fail[l] = 1 when a pivot is not positive
// or, for the first n rows, not above eig_tol (an overlap eigenvalue at or below
// the tolerance forces such a pivot).  LDS: ld x 17 doubles.
__global__ void __launch_bounds__(kCT) k_cholesky(const int* act, const int* nvec, int ld, double* const* Sp,
                                                  double* const* Lp, double eig_tol, int* fail) {
    extern __shared__ double P[];
    const int l = blockIdx.x;
    const int n = nvec[act[l]];
    gptr<const double> S = (gptr<const double>)Sp[l];
    gptr<double> L = (gptr<double>)Lp[l];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, li = lane & 15, lg = lane >> 4;
    int bad = 0;
    for (size_t q = tid; q < (size_t)ld * ld; q += kCT) {
        const int r = (int)(q % ld), c = (int)(q / ld);
        L[q] = r >= c ? S[q] : 0.0;
    }
    __syncthreads();
    for (int k0 = 0; k0 < ld; k0 += kCB) {
        const int m = ld - k0;  // panel rows k0 .. ld-1
        for (int q = tid; q < m * kCB; q += kCT) {
            const int r = q % m, c = q / m;
            P[r * kPS + c] = r >= c ? L[(k0 + r) + (size_t)(k0 + c) * ld] : 0.0;
        }
        __syncthreads();
        // Column j of the panel stays unscaled while the later columns are
        // updated (a_rc -= a_rj a_cj / a_jj), so a step only writes columns
        // > j and reads column j: one barrier per column.
        for (int j = 0; j < kCB; ++j) {
            const double piv = P[j * kPS + j];
            const bool ok = piv > 0.0;  // false for NaN
            if (tid == 0 && (!ok || (k0 + j < n && !(piv > eig_tol)))) bad = 1;
            const double ip = ok ? 1.0 / piv : 0.0;
            for (int r = j + 1 + tid; r < m; r += kCT) {
                const double f = P[r * kPS + j] * ip;
#pragma unroll
                for (int c = 1; c < kCB; ++c)
                    if (c > j && c <= r) P[r * kPS + c] -= f * P[c * kPS + j];
            }
            __syncthreads();
        }
        // L_jj = sqrt(a_jj), L_rj = a_rj / L_jj (the diagonal of P keeps a_jj)
        for (int q = tid; q < m * kCB; q += kCT) {
            const int r = q % m, c = q / m;
            const double piv = P[c * kPS + c];
            const double d = piv > 0.0 ? sqrt(piv) : 1.0;
            double v = 0.0;
            if (r == c) v = d;
            else if (r > c) {
                v = P[r * kPS + c] / d;
                P[r * kPS + c] = v;
            }
            L[(k0 + r) + (size_t)(k0 + c) * ld] = v;
        }
        __syncthreads();
        // trailing update L22 -= L21 L21^T on 16 x 16 tiles (ti >= tj), computed
        // transposed so that a lane's element e is row l%16 of column 4e + l/16
        // (16 consecutive rows per column: coalesced)
        const int T = (m - kCB) / 16;
        const int ntile = T * (T + 1) / 2;
        for (int q = wid; q < ntile; q += kCW) {
            int ti = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
            while ((ti + 1) * (ti + 2) / 2 <= q) ++ti;
            while (ti * (ti + 1) / 2 > q) --ti;
            const int tj = q - ti * (ti + 1) / 2;
            const int pr = kCB + 16 * ti, pc = kCB + 16 * tj;  // panel rows of the tile's rows / columns
            const int gr = k0 + pr, gc = k0 + pc;
            dvec4 acc;
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[e] = L[(gr + li) + (size_t)(gc + 4 * e + lg) * ld];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const double a = -P[(pc + li) * kPS + 4 * s + lg];  // A[i][k] = -L21[col i][4s + k]
                const double b = P[(pr + li) * kPS + 4 * s + lg];   // B[k][j] = L21[row j][4s + k]
                acc = mfma4(a, b, acc);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = li, col = 4 * e + lg;
                if (ti != tj || row >= col) L[(gr + row) + (size_t)(gc + col) * ld] = acc[e];
            }
        }
        __syncthreads();
    }
    if (tid == 0) fail[l] = bad;
}

// Cb = (L^-1)^T (upper; zero below the diagonal) from the lower factor L.
// LDS: 2 x 16 x 17 doubles.
__global__ void __launch_bounds__(kCT) k_triinv_cb(const int* act, const int* nvec, int ld, double* const* Lp,
                                                   double* const* Cbp) {
    __shared__ double Ld[kCB * kPS], Xd[kCB * kPS];
    const int l = blockIdx.x;
    (void)act;
    (void)nvec;
    gptr<const double> L = (gptr<const double>)Lp[l];
    gptr<double> Cb = (gptr<double>)Cbp[l];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, li = lane & 15, lg = lane >> 4;
    const int nb = ld / kCB;
    // blocks strictly below the block diagonal of Cb are zero
    for (size_t q = tid; q < (size_t)ld * ld; q += kCT) {
        const int r = (int)(q % ld), c = (int)(q / ld);
        if (r / kCB > c / kCB) Cb[q] = 0.0;
    }
    for (int i = 0; i < nb; ++i) {
        const int i0 = kCB * i;
        if (tid < kCB * kCB) {
            const int r = tid % kCB, c = tid / kCB;
            Ld[r * kPS + c] = L[(i0 + r) + (size_t)(i0 + c) * ld];
        }
        __syncthreads();
        // X_ii = L_ii^-1, column c by forward substitution
        if (tid < kCB) {
            const int c = tid;
            double x[kCB];
#pragma unroll
            for (int r = 0; r < kCB; ++r) {
                double s = r == c ? 1.0 : 0.0;
#pragma unroll
                for (int q = 0; q < kCB; ++q)
                    if (q < r) s -= Ld[r * kPS + q] * x[q];
                x[r] = r < c ? 0.0 : s / Ld[r * kPS + r];
            }
#pragma unroll
            for (int r = 0; r < kCB; ++r) Xd[r * kPS + c] = x[r];
        }
        __syncthreads();
        if (tid < kCB * kCB) {
            const int r = tid / kCB, c = tid % kCB;
            Cb[(i0 + c) + (size_t)(i0 + r) * ld] = Xd[r * kPS + c];
        }
        // X_ij = -X_ii T, T = sum_{k=j}^{i-1} L_ik X_kj, one block column per wavefront
        for (int j = wid; j < i; j += kCW) {
            const int j0 = kCB * j;
            dvec4 t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
            for (int kq = j0; kq < i0; kq += 4) {
                const double a = L[(i0 + li) + (size_t)(kq + lg) * ld];   // A[i'][k] = L[i0 + i'][kq + k]
                const double b = Cb[(j0 + li) + (size_t)(kq + lg) * ld];  // B[k][j'] = X[kq + k][j0 + j']
                t = mfma4(a, b, t);
            }
            // k-step s of X_ii T takes element s of t as its B operand
            dvec4 rr = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < 4; ++s) rr = mfma4(Xd[li * kPS + 4 * s + lg], t[s], rr);
#pragma unroll
            for (int e = 0; e < 4; ++e) Cb[(j0 + li) + (size_t)(i0 + 4 * e + lg) * ld] = -rr[e];
        }
        __syncthreads();
    }
}

hipError_t launch_cholesky_cb(hipStream_t s, int count, const int* act, const int* nvec, int ld, double* const* S,
                              double* const* L, double* const* Cb, double eig_tol, int* fail) {
    if (ld > kCholMaxDim || ld % 32 != 0) return hipErrorInvalidValue;
    if (count == 0) return hipSuccess;
    const size_t lds = (size_t)ld * kPS * sizeof(double);
    static bool attr = false;  // LDS above 64 KiB must be requested explicitly
    if (!attr) {
        hipError_t err = hipFuncSetAttribute((const void*)k_cholesky, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)((size_t)kCholMaxDim * kPS * sizeof(double)));
        if (err != hipSuccess) return err;
        attr = true;
    }
    hipLaunchKernelGGL(k_cholesky, dim3(count), dim3(kCT), lds, s, act, nvec, ld, S, L, eig_tol, fail);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_triinv_cb, dim3(count), dim3(kCT), 0, s, act, nvec, ld, (double* const*)L, Cb);
    return hipGetLastError();
}

}  // namespace gfn1
