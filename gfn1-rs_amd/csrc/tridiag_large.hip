// Householder tridiagonalisation of one large symmetric matrix over the whole
// device: LAPACK dsytrd with uplo = 'L' (dlatrd panels of kLargeTridiagPanel
// reflectors, the trailing matrix updated once per panel on the matrix cores).
//
// Per column i of a panel (position j), three kernels:
//   k_lt_column  one work-group: finishes W(:, j-1) = p + a2 v, forms column i
//                from the partial update k_lt_reduce left, and generates the
//                reflector (dlarfg) into A, V(:, j), d, e, tau;
//   k_lt_symv    one work-group per 64 x 64 tile of the trailing lower triangle:
//                y = A v, the row part and the column part of each tile written
//                to their own slot of `part` (every slot written exactly once, so
//                the sum is deterministic), plus the dot products V^T v, W^T v;
//   k_lt_reduce  one work-group per 256 rows: p = tau (y - V W^T v - W V^T v)
//                into W(:, j), partial sums of p.v, and the next column after all
//                of the panel's corrections but the one that needs p.v.
// The symmetric product reads the trailing lower triangle once per column
// (8 (n-i)^2 / 2 bytes), across every CU.
#include "tridiag_large.hpp"
#include "wave_reduce.hpp"

#include <algorithm>
#include <cmath>

namespace gfn1 {
namespace {

constexpr int kNB = kLargeTridiagPanel;
static_assert(kNB == 32, "the dot products run as two groups of 16 panel columns");
constexpr int kTile = 64;
constexpr int kChunk = 256;  // rows per work-group of k_lt_reduce and of the dot products
constexpr int kColThreads = 1024;
constexpr int kRowsPerThread = kLargeTridiagMaxDim / kColThreads;
typedef double tvec4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Sum over the work-group in a fixed order; every thread gets lane 0's sums.
template <int NT>
__device__ double group_sum(double v, double* red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < NT / 64; ++k) t += red[k];
    return t;
}

__global__ void __launch_bounds__(kColThreads)
    k_lt_column(int n, int ld, int i, int j, int fresh, int finalize_only, double* __restrict__ A,
                double* __restrict__ V, double* __restrict__ W, int ldv, const double* __restrict__ next,
                const double* __restrict__ pv, int npv, double* __restrict__ d, double* __restrict__ e,
                double* __restrict__ tau) {
    __shared__ double red[kColThreads / 64];
    __shared__ double sh[2];
    const int tid = threadIdx.x;
    double s = 0.0;
    if (!fresh) {
        // W(:, j-1) = p + a2 v with a2 = -tau_{i-1} (p . v) / 2 (rows >= i)
        // npv <= 64: every wavefront forms the same sum (same lanes, same order)
        const int l = tid & 63;
        const double pvs = wave_sum(l < npv ? pv[l] : 0.0);
        const double a2 = -0.5 * tau[i - 1] * pvs;
        const double* vp = V + (size_t)(j - 1) * ldv;
        double* wp = W + (size_t)(j - 1) * ldv;
        // the part of column i's update that needs a2: v_r (p_i + 2 a2 v_i)
        s = wp[i] + 2.0 * a2 * vp[i];
        __syncthreads();  // every thread has read wp[i] before it is rewritten
        for (int r = i + tid; r < n; r += kColThreads) wp[r] += a2 * vp[r];
    }
    if (finalize_only) return;
    double a[kRowsPerThread];
    const double* vp = fresh ? nullptr : V + (size_t)(j - 1) * ldv;
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
        const int r = i + tid + k * kColThreads;
        a[k] = 0.0;
        if (r < n) a[k] = fresh ? A[r + (size_t)i * ld] : next[r] - vp[r] * s;
    }
    if (tid == 0) sh[0] = a[0];
    if (tid == 1) sh[1] = a[0];
    double xn = 0.0;
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
        const int r = i + tid + k * kColThreads;
        if (r >= i + 2 && r < n) xn += a[k] * a[k];
    }
    const double xnorm2 = group_sum<kColThreads>(xn, red);
    const double dd = sh[0], alpha = sh[1];
    double beta = alpha, tk = 0.0, scale = 0.0;
    if (xnorm2 > 0.0) {
        const double nr = sqrt(alpha * alpha + xnorm2);
        beta = alpha >= 0.0 ? -nr : nr;
        tk = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
    }
    double* vj = V + (size_t)j * ldv;
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
        const int r = i + tid + k * kColThreads;
        if (r >= n) continue;
        if (r == i) {
            A[r + (size_t)i * ld] = dd;
        } else if (r == i + 1) {
            A[r + (size_t)i * ld] = beta;
            vj[r] = 1.0;
        } else {
            const double v = a[k] * scale;
            A[r + (size_t)i * ld] = v;
            vj[r] = v;
        }
    }
    if (tid == 0) {
        d[i] = dd;
        e[i] = beta;
        tau[i] = tk;
    }
}

// Strips of the trailing lower triangle (rows and columns >= c0): tile row RB
// (64 rows) is split into strips of up to `strip` 64 x 64 tiles, column blocks
// CB0 .. CB0 + strip - 1 (at most RB).  Strips are numbered row by row;
// strips_before(rho, strip) counts those of the tile rows above row rho.
__host__ __device__ __forceinline__ int strips_before(int rho, int strip) {
    const int k = rho / strip;
    return rho + strip * k * (k - 1) / 2 + k * (rho - k * strip);
}
// Strip length of a column with nt tile rows: long strips while there are
// tiles enough to give every CU several work-groups, shorter ones below.
int strip_length(int nt) {
    for (int strip = 8; strip > 1; strip >>= 1)
        if (strips_before(nt, strip) >= 2048) return strip;
    return 1;
}

// y = A v over the trailing lower triangle.  A work-group takes one strip:
// lane l holds row R0 + l, wavefront w columns 16 w .. 16 w + 15 of each
// tile, and the next tile's 16 loads are in flight while the current one is
// used.  The row part (row r times the columns of the strip) is summed over
// the strip in registers; the column part of each tile (column c times the
// rows below the diagonal) by a transposed wavefront sum.  Slots: part[b][r]
// for b <= block(r) come from the row's strips (the strip total in its first
// block, zero in the others, plus the diagonal tile's column part in block(r)
// itself); part[b][r] for b > block(r) is the column part of tile (b,
// block(r)).  Every slot is written exactly once per column, so the sum over
// b is deterministic.  Work-groups from nstrips on form the dot products V^T
// v, W^T v over 256-row chunks.
__global__ void __launch_bounds__(256)
    k_lt_symv(int n, int ld, int c0, int j, const double* __restrict__ A, const double* __restrict__ V,
              const double* __restrict__ W, int ldv, double* __restrict__ part, int ldp, double* __restrict__ dots,
              int nstrips, int nt, int strip) {
    __shared__ double prs[4][kTile];
    __shared__ double pdc[kTile];
    __shared__ double wdots[4][2 * kNB];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const double* v = V + (size_t)j * ldv;
    const int t = blockIdx.x;
    if (t >= nstrips) {
        // V^T v and W^T v of a 256-row chunk: 16 products per lane summed over
        // the wavefront by one transposed reduction, the four wavefronts in LDS
        const int ch = t - nstrips;
        const int r = c0 + ch * kChunk + tid;
        const double vv = r < n ? v[r] : 0.0;
#pragma unroll
        for (int g = 0; g < 2 * kNB / 16; ++g) {
            const int jj0 = (g & 1) * 16;
            if (jj0 >= j) continue;  // uniform: depends on j only
            const double* M = g < 2 ? V : W;
            double cp[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) cp[q] = (r < n && jj0 + q < j) ? M[(size_t)(jj0 + q) * ldv + r] * vv : 0.0;
            const double tot = wave::reduce16(cp, lane);
            if (lane < 16) wdots[w][g * 16 + wave::reduce16_index(lane)] = tot;
        }
        __syncthreads();
        if (tid < 2 * kNB) {
            const int jj = tid & (kNB - 1);
            if (jj < j)
                dots[(size_t)ch * 2 * kNB + tid] = ((wdots[0][tid] + wdots[1][tid]) + wdots[2][tid]) + wdots[3][tid];
        }
        return;
    }
    int lo = 0, hi = nt - 1;  // the tile row of strip t: largest rho with strips_before(rho, strip) <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (strips_before(mid, strip) <= t) lo = mid;
        else hi = mid - 1;
    }
    const int b0 = c0 / kTile, rho = lo;
    const int RB = b0 + rho, CB0 = b0 + (t - strips_before(rho, strip)) * strip;
    const int CBe = min(CB0 + strip - 1, RB);
    const int R0 = RB * kTile, r = R0 + lane;
    const bool rok = r < n && r >= c0;
    const double vr = rok ? v[r] : 0.0;
    auto load = [&](double (&x)[16], int cb) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = cb * kTile + 16 * w + q;
            const bool ok = rok && c < n && c >= c0 && r >= c;
            x[q] = ok ? A[r + (size_t)c * ld] : 0.0;
        }
    };
    double x[16], xn[16];
    load(x, CB0);
    double rs = 0.0;
    for (int cb = CB0; cb <= CBe; ++cb) {
        if (cb < CBe) load(xn, cb + 1);
        const int C0 = cb * kTile + 16 * w;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = C0 + q;
            const double vc = (c < n && c >= c0) ? v[c] : 0.0;
            rs += x[q] * vc;
            x[q] = (cb != RB || r > c) ? x[q] * vr : 0.0;  // the column product, in place
        }
        const double tot = wave::reduce16(x, lane);
        const int cq = C0 + wave::reduce16_index(lane);
        if (lane < 16 && cq < n) {
            if (cb != RB) part[(size_t)RB * ldp + cq] = tot;
            else pdc[cq - R0] = tot;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) x[q] = xn[q];
    }
    prs[w][lane] = rs;
    __syncthreads();
    if (tid < kTile) {
        const int rr = R0 + tid;
        if (rr < n) {
            const double rowtot = prs[0][tid] + prs[1][tid] + prs[2][tid] + prs[3][tid];
            for (int b = CB0; b <= CBe; ++b)
                part[(size_t)b * ldp + rr] = (b == CB0 ? rowtot : 0.0) + (b == RB ? pdc[tid] : 0.0);
        }
    }
}

__global__ void __launch_bounds__(kChunk)
    k_lt_reduce(int n, int ld, int c0, int j, int nblk, const double* __restrict__ part, int ldp,
                const double* __restrict__ A, const double* __restrict__ V, double* __restrict__ W, int ldv,
                const double* __restrict__ dots, int ndot, const double* __restrict__ tau, double* __restrict__ pv,
                double* __restrict__ next, int last_in_panel) {
    __shared__ double dv[kNB], dw[kNB], wi[kNB], vi[kNB];
    __shared__ double s4[4][2 * kNB];
    __shared__ double red[kChunk / 64];
    const int tid = threadIdx.x;
    {
        // the chunk sums of V^T v, W^T v: four interleaved partial sums per value
        const int idx = tid & (2 * kNB - 1), q = tid >> 6;
        double acc = 0.0;
        if ((idx & (kNB - 1)) < j)
            for (int ch = q; ch < ndot; ch += 4) acc += dots[(size_t)ch * 2 * kNB + idx];
        s4[q][idx] = acc;
    }
    if (tid < j) {
        wi[tid] = W[(size_t)tid * ldv + c0];
        vi[tid] = V[(size_t)tid * ldv + c0];
    }
    __syncthreads();
    if (tid < 2 * kNB && (tid & (kNB - 1)) < j) {
        const double tot = ((s4[0][tid] + s4[1][tid]) + s4[2][tid]) + s4[3][tid];
        if (tid < kNB) dv[tid] = tot;  // V^T v
        else dw[tid - kNB] = tot;      // W^T v
    }
    __syncthreads();
    const int b0 = c0 / kTile;
    const double tk = tau[c0 - 1];
    const double* v = V + (size_t)j * ldv;
    const double vc0 = v[c0];
    const int r = c0 + blockIdx.x * kChunk + tid;
    double pvl = 0.0;
    if (r < n) {
        // eight independent sums keep eight loads in flight; fixed combination order
        double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        int b = b0;
        for (; b + 8 <= nblk; b += 8) {
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[u] += part[(size_t)(b + u) * ldp + r];
        }
        for (int u = 0; b < nblk; ++b, ++u) acc[u] += part[(size_t)b * ldp + r];
        const double y = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
        double corr = 0.0, upd = 0.0;
#pragma unroll 4
        for (int jj = 0; jj < j; ++jj) {
            const double vr = V[(size_t)jj * ldv + r], wr = W[(size_t)jj * ldv + r];
            corr += vr * dw[jj] + wr * dv[jj];
            upd += vr * wi[jj] + wr * vi[jj];
        }
        const double p = tk * (y - corr);
        W[(size_t)j * ldv + r] = p;
        pvl = p * v[r];
        if (!last_in_panel) next[r] = A[r + (size_t)c0 * ld] - upd - p * vc0;
    }
    const double sum = group_sum<kChunk>(pvl, red);
    if (tid == 0) pv[blockIdx.x] = sum;
}

// A(c:, c:) -= V W^T + W V^T on the lower triangle of the trailing block (m =
// n - c rows; A, V, W offset to row and column c), on the FP64 matrix cores.
// One work-group per 64 x 64 tile on or below the diagonal; wavefront w takes
// rows 16 w .. 16 w + 15 of the tile and its four 16 x 16 column blocks.  The
// block sits in the accumulator transposed (element e of lane (li, lg) is
// A(R0 + li, C0 + 4 e + lg)), so its loads and stores are column-contiguous;
// D' += Aop Bop over K = 2 kNB with Aop = -[W_C | V_C] and Bop = [V_R | W_R]^T.
__global__ void __launch_bounds__(256)
    k_lt_syr2k(int m, int jb, double* __restrict__ A, int ld, const double* __restrict__ V,
               const double* __restrict__ W, int ldv) {
    const int t = blockIdx.x;
    int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= t) ++I;
    while (I * (I + 1) / 2 > t) --I;
    const int J = t - I * (I + 1) / 2;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lg = lane >> 4;
    const int R0 = I * kTile + 16 * w;
    if (R0 >= m) return;  // no barrier below: wavefronts are independent
    const int rB = R0 + li;
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        const int C0 = J * kTile + 16 * q;
        if (C0 >= m || C0 > R0 + 15) continue;  // outside the matrix or above the diagonal
        const int cA = C0 + li;
        tvec4 acc;
#pragma unroll
        for (int e2 = 0; e2 < 4; ++e2) {
            const int r = R0 + li, c = C0 + 4 * e2 + lg;
            acc[e2] = (r < m && c < m && r >= c) ? A[r + (size_t)c * ld] : 0.0;
        }
#pragma unroll
        for (int k0 = 0; k0 < 2 * kNB; k0 += 4) {
            const int kk = k0 + lg, kc = kk & (kNB - 1);
            const bool upper = kk >= kNB, ok = kc < jb;
            const double a = (ok && cA < m) ? -(upper ? V : W)[(size_t)kc * ldv + cA] : 0.0;
            const double b = (ok && rB < m) ? (upper ? W : V)[(size_t)kc * ldv + rB] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int e2 = 0; e2 < 4; ++e2) {
            const int r = R0 + li, c = C0 + 4 * e2 + lg;
            if (r < m && c < m && r >= c) A[r + (size_t)c * ld] = acc[e2];
        }
    }
}

__global__ void k_lt_last(int n, int ld, const double* __restrict__ A, double* d, double* e, double* tau) {
    d[n - 1] = A[(size_t)(n - 1) * ld + (n - 1)];
    e[n - 1] = 0.0;
    tau[n - 1] = 0.0;
}

}  // namespace

hipError_t LargeTridiagWork::allocate(int rows) {
    if (rows <= n) return hipSuccess;
    release();
    const size_t nblk = ((size_t)rows + kTile - 1) / kTile, nch = ((size_t)rows + kChunk - 1) / kChunk;
    hipError_t err;
    if ((err = hipMalloc(&V, sizeof(double) * (size_t)rows * kNB)) != hipSuccess) return err;
    if ((err = hipMalloc(&W, sizeof(double) * (size_t)rows * kNB)) != hipSuccess) return err;
    if ((err = hipMalloc(&part, sizeof(double) * nblk * rows)) != hipSuccess) return err;
    if ((err = hipMalloc(&dots, sizeof(double) * nch * 2 * kNB)) != hipSuccess) return err;
    if ((err = hipMalloc(&pv, sizeof(double) * nch)) != hipSuccess) return err;
    if ((err = hipMalloc(&next, sizeof(double) * rows)) != hipSuccess) return err;
    n = rows;
    return hipSuccess;
}

void LargeTridiagWork::release() {
    for (double** p : {&V, &W, &part, &dots, &pv, &next}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    n = 0;
}

hipError_t launch_tridiag_large(hipStream_t s, int n, int ld, double* A, double* d, double* e, double* tau,
                                LargeTridiagWork& w) {
    if (n < 2 || n > kLargeTridiagMaxDim || ld < n || w.n < n) return hipErrorInvalidValue;
    const int ldv = w.n, ldp = w.n;
    const int nblk = (n + kTile - 1) / kTile;
    int npv = 0;
    for (int i0 = 0; i0 < n - 1; i0 += kNB) {
        const int jb = std::min(kNB, n - 1 - i0);
        for (int j = 0; j < jb; ++j) {
            const int i = i0 + j, c0 = i + 1;
            hipLaunchKernelGGL(k_lt_column, dim3(1), dim3(kColThreads), 0, s, n, ld, i, j, j == 0 ? 1 : 0, 0, A,
                               w.V, w.W, ldv, (const double*)w.next, (const double*)w.pv, npv, d, e, tau);
            const int b0 = c0 / kTile, nt = nblk - b0;
            const int strip = strip_length(nt);
            const int nstrips = strips_before(nt, strip);
            const int nch = (n - c0 + kChunk - 1) / kChunk;
            const int ndot = j > 0 ? nch : 0;
            hipLaunchKernelGGL(k_lt_symv, dim3(nstrips + ndot), dim3(256), 0, s, n, ld, c0, j, (const double*)A,
                               (const double*)w.V, (const double*)w.W, ldv, w.part, ldp, w.dots, nstrips, nt, strip);
            hipLaunchKernelGGL(k_lt_reduce, dim3(nch), dim3(kChunk), 0, s, n, ld, c0, j, nblk,
                               (const double*)w.part, ldp, (const double*)A, (const double*)w.V, w.W, ldv,
                               (const double*)w.dots, ndot, (const double*)tau, w.pv, w.next, j == jb - 1 ? 1 : 0);
            npv = nch;
        }
        // the panel's last W column, then A(c:, c:) -= V W^T + W V^T
        const int c = i0 + jb;
        hipLaunchKernelGGL(k_lt_column, dim3(1), dim3(kColThreads), 0, s, n, ld, c, jb, 0, 1, A, w.V, w.W, ldv,
                           (const double*)w.next, (const double*)w.pv, npv, d, e, tau);
        const int m = n - c, mt = (m + kTile - 1) / kTile;
        hipLaunchKernelGGL(k_lt_syr2k, dim3(mt * (mt + 1) / 2), dim3(256), 0, s, m, jb, A + c + (size_t)c * ld, ld,
                           (const double*)(w.V + c), (const double*)(w.W + c), ldv);
    }
    hipLaunchKernelGGL(k_lt_last, dim3(1), dim3(1), 0, s, n, ld, (const double*)A, d, e, tau);
    return hipGetLastError();
}

}  // namespace gfn1
