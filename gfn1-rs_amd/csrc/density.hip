// Occupied orbitals and density matrices of the SCC on the FP64 matrix cores.
//
//   k_orbitals   C = Cb Z: the orbitals of the occupied (and partially occupied)
//                eigenvectors Z of the reduced problem, Cb = L^-T upper
//                triangular, so each 64 x 64 tile of C sums only over the rows
//                of Z at or below its first row
//   k_density    M = C diag(w) C^T with w = f (density P) or eps f (energy-
//                weighted density W), column_weighted_gram (linalg.rs:243-273).
//                Only the tiles on and below the block diagonal are computed;
//                each is written to both triangles, so M is exactly symmetric
//                (the diagonal tiles are symmetrised as (M + M^T)/2, as the
//                reference symmetrises).  For P the epilogue also forms the
//                Mulliken row sums sum_nu S_mu,nu P_mu,nu and sum_nu H0 P
//                (electronic.rs:3160-3188) per 64-column block: a tile below
//                the diagonal contributes row sums to its row block and column
//                sums to its column block, so every (block, row) partial is
//                written exactly once and the reduction order is fixed.
//
// Both kernels use 64 x 64 output tiles on 256 threads: every wavefront owns a
// 32 x 32 quarter (2 x 2 tiles of v_mfma_f64_16x16x4f64), the K dimension runs
// in chunks of 16 staged through LDS as [k][row] slabs, the next chunk's
// operands are loaded into registers while the current one is multiplied, and
// the output tile goes through LDS so that global writes are column-contiguous.
// The work-groups of one matrix are placed on one XCD (dispatch order is
// round-robin over the eight XCDs), so the operands they share stay in that
// XCD's L2.
#include <hip/hip_runtime.h>

#include "density.hpp"

namespace gfn1 {
namespace {

typedef double dvec4 __attribute__((ext_vector_type(4)));

constexpr int kT = 64;           // output tile edge
constexpr int kK = 16;           // K chunk
constexpr int kThreads = 256;
constexpr int kSlab = kT + 4;    // LDS stride of a [k][row] slab (doubles)
constexpr int kOS = kT + 1;      // LDS stride of the output tile, [col][row] (odd: conflict-free transposed reads)
constexpr int kXcds = 8;

__device__ __forceinline__ dvec4 mfma(double a, double b, dvec4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// consecutive logical work-group ids on one XCD (gridDim.x is a multiple of 8)
__device__ __forceinline__ int logical_block() {
    const int id = blockIdx.x, per = gridDim.x / kXcds;
    return (id % kXcds) * per + id / kXcds;
}

// Deterministic wavefront sum (every lane gets the total).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One K chunk of the 64 x 64 tile: acc[rt][ct] += A(rows) B(cols) over the 16
// staged k values.  The A operand of k-step s is slab A[4 s + lg][row], the B
// operand slab B[4 s + lg][col] (lane li = row or column within a 16-tile).
__device__ __forceinline__ void tile_chunk(const double* As, const double* Bs, int r0, int c0, int li, int lg,
                                           dvec4 (&acc)[2][2]) {
#pragma unroll
    for (int s = 0; s < kK / 4; ++s) {
        double a[2], b[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            a[t] = As[(4 * s + lg) * kSlab + r0 + 16 * t + li];
            b[t] = Bs[(4 * s + lg) * kSlab + c0 + 16 * t + li];
        }
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = mfma(a[rt], b[ct], acc[rt][ct]);
    }
}

// Accumulators to the output tile in LDS: element e of lane (li, lg) of tile
// (rt, ct) is row r0 + 16 rt + 4 e + lg, column c0 + 16 ct + li.
__device__ __forceinline__ void tile_to_lds(const dvec4 (&acc)[2][2], double* Os, int r0, int c0, int li, int lg) {
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                Os[(c0 + 16 * ct + li) * kOS + r0 + 16 * rt + 4 * e + lg] = acc[rt][ct][e];
}

}  // namespace

// C[l] (ld x kcap, column-major) = Cb[l] Z[l] with Z element (m, j) at
// Z[m * zr + j * zc]; rows m >= n of Z and columns j >= kneed are zero, so C is
// zero there too.  Grid: count * ntr * ntc work-groups (rounded up to a
// multiple of 8), ntr = ld / 64 rounded up, ntc = kcap / 64 rounded up.
__global__ void __launch_bounds__(kThreads) k_orbitals(const int* __restrict__ act, const int* __restrict__ nvec,
                                                       const SccState* __restrict__ st, int ld, int kcap,
                                                       double* const* Cbp, double* const* Zp, int zr, int zc,
                                                       double* const* Cp, int count, int ntr, int ntc) {
    __shared__ double As[kK * kSlab], Bs[kK * kSlab], Os[kT * kOS];
    const int logical = logical_block();
    const int per = ntr * ntc;
    const int l = logical / per;
    if (l >= count) return;
    const int I = (logical % per) / ntc, J = (logical % per) % ntc;
    const int inst = act[l];
    if (inst < 0) return;  // gated off in a captured iteration
    const int n = nvec[inst];
    const int kn = min(st[inst].kneed, kcap);
    const int I0 = I * kT, J0 = J * kT;
    if (I0 >= ld || J0 >= kcap) return;
    gptr<const double> Cb = (gptr<const double>)Cbp[l];
    gptr<const double> Z = (gptr<const double>)Zp[l];
    gptr<double> C = (gptr<double>)Cp[l];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, li = lane & 15, lg = lane >> 4;
    const int r0 = (wid >> 1) * 32, c0 = (wid & 1) * 32;
    dvec4 acc[2][2];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = dvec4{0.0, 0.0, 0.0, 0.0};
    // Cb is upper triangular: only rows m >= I0 of Z contribute; Z has no rows >= n
    const int mlo = I0, mhi = J0 < kn ? n : mlo;
    // staging: thread t loads slab element (k = t / 64 + 4 q, row/col = t % 64), q = 0..3
    const int sr = tid & 63, sk = tid >> 6;
    double pa[4], pb[4];
    auto load = [&](int m0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int m = m0 + sk + 4 * q;
            const bool mok = m < mhi;
            pa[q] = (mok && I0 + sr < ld) ? Cb[(size_t)m * ld + I0 + sr] : 0.0;
            pb[q] = (mok && J0 + sr < kcap) ? Z[(size_t)m * zr + (size_t)(J0 + sr) * zc] : 0.0;
        }
    };
    if (mlo < mhi) load(mlo);
    for (int m0 = mlo; m0 < mhi; m0 += kK) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            As[(sk + 4 * q) * kSlab + sr] = pa[q];
            Bs[(sk + 4 * q) * kSlab + sr] = pb[q];
        }
        __syncthreads();
        if (m0 + kK < mhi) load(m0 + kK);
        tile_chunk(As, Bs, r0, c0, li, lg, acc);
    }
    tile_to_lds(acc, Os, r0, c0, li, lg);
    __syncthreads();
    // column-contiguous writes of the tile (rows < ld, columns < kcap)
    for (int q = tid; q < kT * kT; q += kThreads) {
        const int r = q & 63, c = q >> 6;
        if (I0 + r < ld && J0 + c < kcap) C[(size_t)(J0 + c) * ld + I0 + r] = Os[c * kOS + r];
    }
}

// M[l] = C diag(w) C^T on the n x n block (w_j = occ_j, or eps_j occ_j when
// weighted), j < min(kneed, kcap).  Grid: count * nt (nt + 1) / 2 work-groups
// (rounded up to a multiple of 8), nt = ld / 64 rounded up; tile t of a matrix
// is the t-th (I, J) pair with J <= I in row order.  With rowS / rowH (P only):
// rowS[blk * stride + ao0 + mu] = sum over nu in column block blk of S P,
// likewise rowH with H0, for every block blk < ceil(n / 64).
__global__ void __launch_bounds__(kThreads) k_density(const int* __restrict__ act, const int* __restrict__ nvec,
                                                      const SccState* __restrict__ st, int ld, int kcap,
                                                      double* const* Cp, const double* __restrict__ eps_all,
                                                      const double* __restrict__ occ_all,
                                                      const int64_t* __restrict__ vec0, int weighted,
                                                      double* const* Mp, double* const* Sp, double* const* Hp,
                                                      const int* __restrict__ ao0, double* rowS, double* rowH,
                                                      int stride, int count, int nt) {
    __shared__ double As[kK * kSlab], Bs[kK * kSlab], Os[kT * kOS], red[2][4][kT];
    const int logical = logical_block();
    const int per = nt * (nt + 1) / 2;
    const int l = logical / per;
    if (l >= count) return;
    int t = logical % per, I = 0;
    while (t > I) {
        t -= I + 1;
        ++I;
    }
    const int J = t;
    const int inst = act[l];
    if (inst < 0) return;  // gated off in a captured iteration
    const int n = nvec[inst];
    const int kn = min(st[inst].kneed, kcap);
    const int I0 = I * kT, J0 = J * kT;
    if (I0 >= n) return;
    gptr<const double> C = (gptr<const double>)Cp[l];
    gptr<double> M = (gptr<double>)Mp[l];
    const double* occ = occ_all + vec0[inst];
    const double* eps = eps_all + vec0[inst];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, li = lane & 15, lg = lane >> 4;
    const int r0 = (wid >> 1) * 32, c0 = (wid & 1) * 32;
    dvec4 acc[2][2];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = dvec4{0.0, 0.0, 0.0, 0.0};
    const int sr = tid & 63, sk = tid >> 6;
    double pa[4], pb[4];
    auto load = [&](int j0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = j0 + sk + 4 * q;
            const bool jok = j < kn;
            const double w = jok ? (weighted ? eps[j] * occ[j] : occ[j]) : 0.0;
            pa[q] = (jok && I0 + sr < n) ? C[(size_t)j * ld + I0 + sr] * w : 0.0;
            pb[q] = (jok && J0 + sr < n) ? C[(size_t)j * ld + J0 + sr] : 0.0;
        }
    };
    if (kn > 0) load(0);
    for (int j0 = 0; j0 < kn; j0 += kK) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            As[(sk + 4 * q) * kSlab + sr] = pa[q];
            Bs[(sk + 4 * q) * kSlab + sr] = pb[q];
        }
        __syncthreads();
        if (j0 + kK < kn) load(j0 + kK);
        tile_chunk(As, Bs, r0, c0, li, lg, acc);
    }
    tile_to_lds(acc, Os, r0, c0, li, lg);
    __syncthreads();
    const bool diag = I == J;
    if (diag) {
        // (M + M^T) / 2 inside the diagonal tile; both halves read before any write
        double v[kT * kT / kThreads];
#pragma unroll
        for (int u = 0; u < kT * kT / kThreads; ++u) {
            const int q = tid + u * kThreads, r = q & 63, c = q >> 6;
            v[u] = 0.5 * (Os[c * kOS + r] + Os[r * kOS + c]);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kT * kT / kThreads; ++u) {
            const int q = tid + u * kThreads, r = q & 63, c = q >> 6;
            Os[c * kOS + r] = v[u];
        }
        __syncthreads();
    }
    const int rows = min(kT, n - I0), cols = min(kT, n - J0);
    for (int q = tid; q < kT * kT; q += kThreads) {
        const int r = q & 63, c = q >> 6;
        if (r < rows && c < cols) M[(size_t)(J0 + c) * ld + I0 + r] = Os[c * kOS + r];
    }
    if (!diag) {
        // mirror: M(J0 + a, I0 + b) = tile(b, a), contiguous in a
        for (int q = tid; q < kT * kT; q += kThreads) {
            const int a = q & 63, b = q >> 6;
            if (a < cols && b < rows) M[(size_t)(I0 + b) * ld + J0 + a] = Os[a * kOS + b];
        }
    }
    if (!rowS) return;
    // Mulliken sums: wavefront wid takes columns 16 wid .. 16 wid + 15 of the
    // tile, lane = row
    gptr<const double> S = (gptr<const double>)Sp[l];
    gptr<const double> H = (gptr<const double>)Hp[l];
    const int r = lane, base = ao0[inst];
    double ps = 0.0, ph = 0.0;
    for (int cc = 0; cc < 16; ++cc) {
        const int c = 16 * wid + cc;
        double s = 0.0, h = 0.0;
        if (r < rows && c < cols) {
            const double m = Os[c * kOS + r];
            const size_t at = (size_t)(J0 + c) * ld + I0 + r;
            s = S[at] * m;
            h = H[at] * m;
        }
        ps += s;
        ph += h;
        if (!diag) {
            // column sum over the tile's rows: block I's partial of row J0 + c
            const double cs = wave_sum(s), ch = wave_sum(h);
            if (lane == 0 && c < cols) {
                rowS[(size_t)I * stride + base + J0 + c] = cs;
                rowH[(size_t)I * stride + base + J0 + c] = ch;
            }
        }
    }
    red[0][wid][r] = ps;
    red[1][wid][r] = ph;
    __syncthreads();
    if (tid < kT && tid < rows) {
        double s = 0.0, h = 0.0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            s += red[0][w][tid];
            h += red[1][w][tid];
        }
        rowS[(size_t)J * stride + base + I0 + tid] = s;
        rowH[(size_t)J * stride + base + I0 + tid] = h;
    }
}

hipError_t launch_orbitals(hipStream_t s, int count, const int* act, const int* nvec, const SccState* st, int ld,
                           int kcap, double* const* Cb, double* const* Z, int zr, int zc, double* const* C) {
    if (count <= 0) return hipSuccess;
    const int ntr = (ld + kT - 1) / kT, ntc = (kcap + kT - 1) / kT;
    const long long blocks = ((long long)count * ntr * ntc + kXcds - 1) / kXcds * kXcds;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_orbitals, dim3((unsigned)blocks), dim3(kThreads), 0, s, act, nvec, st, ld, kcap, Cb, Z, zr,
                       zc, C, count, ntr, ntc);
    return hipGetLastError();
}

hipError_t launch_density(hipStream_t s, int count, const int* act, const int* nvec, const SccState* st, int ld,
                          int kcap, double* const* C, const double* eps_all, const double* occ_all,
                          const int64_t* vec0, bool weighted, double* const* M, double* const* S,
                          double* const* H0, const int* ao0, double* rowS, double* rowH, int stride) {
    if (count <= 0) return hipSuccess;
    const int nt = (ld + kT - 1) / kT;
    const long long blocks = ((long long)count * nt * (nt + 1) / 2 + kXcds - 1) / kXcds * kXcds;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_density, dim3((unsigned)blocks), dim3(kThreads), 0, s, act, nvec, st, ld, kcap, C, eps_all,
                       occ_all, vec0, weighted ? 1 : 0, M, S, H0, ao0, rowS, rowH, stride, count, nt);
    return hipGetLastError();
}

int density_row_blocks(int ld) { return (ld + kT - 1) / kT; }

}  // namespace gfn1
