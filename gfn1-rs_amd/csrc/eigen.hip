// Batched symmetric eigensolver for the SCC (gfx950).
//
//   k_tridiag         blocked Householder reduction A = Q T Q^T (LAPACK dsytrd
//                     'L' sequence with dlatrd panels), one work-group per
//                     matrix; the symmetric matrix-vector product reads the
//                     lower triangle once per column
//   k_tridiag_eig     all eigenvalues of T by Sturm-count bisection, Fermi
//                     occupations, and the eigenvectors of every orbital with a
//                     non-negligible occupation by inverse iteration (LAPACK
//                     dlagtf/dlagts factorisation), one thread per eigenvector;
//                     groups of nearly equal eigenvalues are re-orthogonalised and
//                     resolved by a Rayleigh-Ritz step on the group
//   k_backtransform   Z <- Q Z in compact-WY blocks of 16 reflectors on the FP64
//                     matrix cores, one work-group per (matrix, block of eigenvectors)
//
// The SCC needs the occupied subspace (density matrix), the energies of all
// orbitals (occupations, orbital-energy output) and, at the end, the
// energy-weighted density; eigenvectors of empty orbitals are never formed.
// Eigenvalues are converged to machine precision, so orbital energies and
// occupations match a full dense eigensolver (the reference uses faer's
// self-adjoint eigensolver, linalg.rs:153-191).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "device.hpp"
#include "eigen.hpp"
#include "wave_reduce.hpp"

namespace gfn1 {
using namespace wave;

// Phase timing (profiling builds only, -DGFN1_PHASE_TIMING): thread 0 of the
// first kPhaseMats work-groups adds the wall-clock ticks (100 MHz) it spends in
// each phase of a kernel to gfn1_phase_ticks[kernel][matrix][phase], read back
// with gfn1_debug_phase_ticks.  Compiled out otherwise.
#ifdef GFN1_PHASE_TIMING
constexpr int kPhaseKernels = 2, kPhaseMats = 64, kPhaseSlots = 8;
__device__ unsigned long long gfn1_phase_ticks[kPhaseKernels][kPhaseMats][kPhaseSlots];
#define PT_INIT()                          \
    unsigned long long pt_last = wall_clock64(); \
    unsigned long long pt_acc[kPhaseSlots] = {};
#define PT_MARK(slot)                                        \
    do {                                                     \
        if (threadIdx.x == 0) {                              \
            const unsigned long long pt_now = wall_clock64(); \
            pt_acc[slot] += pt_now - pt_last;                \
            pt_last = pt_now;                                \
        }                                                    \
    } while (0)
#define PT_STORE(kernel)                                                                   \
    do {                                                                                   \
        if (threadIdx.x == 0 && blockIdx.x < kPhaseMats)                                   \
            for (int s_ = 0; s_ < kPhaseSlots; ++s_) gfn1_phase_ticks[kernel][blockIdx.x][s_] += pt_acc[s_]; \
    } while (0)
#else
#define PT_INIT()
#define PT_MARK(slot)
#define PT_STORE(kernel)
#endif

namespace {

constexpr double kEps = 2.220446049250313e-16;
constexpr double kSafeMin = 2.2250738585072014e-308;
constexpr int kThreads = 256;  // k_tridiag_eig
constexpr int kBT = 16;  // reflectors per compact-WY block in the back-transformation
constexpr int kLB = 8;   // loads issued together in the dependent eigenvector passes


// ---------------------------------------------------------------------------
// Householder tridiagonalisation, one work-group of kTriThreads per matrix
// (n <= kTriThreads).  Thread r owns row r in the row phases and keeps row r
// of the current panel's reflectors V and of W in registers, so inside a
// panel the trailing matrix A - V W^T - W V^T is used without extra memory
// traffic; the rank-2b update is applied once per panel in 64 x 64 tiles with
// V and W staged through LDS.  The matrix-vector product reads the lower
// triangle once, in 64 x 16 tiles shared out over the wavefronts.
typedef double tvec4 __attribute__((ext_vector_type(4)));
constexpr int kTriThreads = 512;
constexpr int kTriWaves = kTriThreads / 64;
constexpr int kTW = 16;  // columns per symv tile (64 x 16): keeps the kernel within 128 VGPRs, two work-groups per CU
constexpr int kNB = kEigPanel;
constexpr int kVS = kNB + 1;  // LDS row stride of the staged panel (odd: no bank conflicts)

__global__ void __launch_bounds__(kTriThreads) k_tridiag(const int* __restrict__ act, const int* __restrict__ nvec,
                                                         int ld, double* const* Aptr, double* dvec, double* evec,
                                                         double* tauvec) {
    extern __shared__ double sm[];
    const int l = blockIdx.x;
    if (act[l] < 0) return;  // gated off in a captured iteration
    const int n = nvec[act[l]];
    gptr<double> A = (gptr<double>)Aptr[l];
    double* d = dvec + (size_t)l * ld;
    double* e = evec + (size_t)l * ld;
    double* tau = tauvec + (size_t)l * ld;
    // column steps
    double* xs = sm;               // column k after the pending panel updates
    double* vs = sm + ld;          // reflector
    double* part = sm + 2 * ld;    // A v, one partial slice per wavefront
    // panel end (same memory): staged V and W rows
    double* Vst = sm;
    double* Wst = sm + (size_t)kVS * ld;
    double* ysp = sm + (size_t)2 * kVS * ld;  // per-wavefront W^T v, V^T v
    double* red = ysp + kTriWaves * 16;       // 16
    double* rowk = red + 16;                  // V[k, :], W[k, :]
    const int tid = threadIdx.x, lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);  // wavefront-uniform
    const int r = tid;
    double vreg[kNB], wreg[kNB];
    PT_INIT();
    for (int k0 = 0; k0 < n - 1; k0 += kNB) {
        const int b = min(kNB, n - 1 - k0);
#pragma unroll
        for (int i = 0; i < kNB; ++i) vreg[i] = wreg[i] = 0.0;
        for (int j = 0; j < b; ++j) {
            const int k = k0 + j, c0 = k + 1;
            // ---- column k of the current matrix, reflector annihilating rows k+2..
            double x = 0.0;
            if (r >= k && r < n) {
                x = A[r + (size_t)k * ld];
#pragma unroll
                for (int i = 0; i < kNB; ++i)
                    if (i < j) x -= vreg[i] * rowk[kNB + i] + wreg[i] * rowk[i];
                xs[r] = x;
            }
            const double xnorm2 = bsum((r > c0 && r < n) ? x * x : 0.0, red);
            const double alpha = xs[c0];
            double beta = alpha, tk = 0.0, scale = 0.0;
            if (xnorm2 > 0.0) {
                const double nr = sqrt(alpha * alpha + xnorm2);
                beta = alpha >= 0.0 ? -nr : nr;
                tk = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            double v = 0.0;
            if (r >= c0 && r < n) {
                v = r == c0 ? 1.0 : x * scale;
                vs[r] = v;
                A[r + (size_t)k * ld] = v;
            }
#pragma unroll
            for (int i = 0; i < kNB; ++i)
                if (i == j) vreg[i] = v;
            if (tid == 0) {
                d[k] = xs[k];
                e[k] = beta;
                tau[k] = tk;
            }
            __syncthreads();
            PT_MARK(0);
            // ---- A_panel v over the lower triangle of the trailing block, in 64 x kTW
            //      tiles (rows c0 + 64t + lane, columns c0 + kTW u .. + kTW - 1, with
            //      u < (64 / kTW)(t + 1)); the row and column contributions of a
            //      wavefront's tiles go to its own slice of part, so the sum is
            //      deterministic
            {
                double* slice = part + wid * ld;
                for (int rr = c0 + lane; rr < n; rr += 64) slice[rr] = 0.0;
                const int mt = (n - c0 + 63) >> 6, mu = (n - c0 + kTW - 1) / kTW;
                int t = 0, u = wid;  // tile q = wid, wid + kTriWaves, ... in row-block order
                for (;;) {
                    while (t < mt && u >= min((64 / kTW) * (t + 1), mu)) {
                        u -= min((64 / kTW) * (t + 1), mu);
                        ++t;
                    }
                    if (t >= mt) break;
                    const int rb = c0 + 64 * t, cb = c0 + kTW * u;  // first row / column (uniform)
                    const int rr = rb + lane;
                    const bool rowok = rr < n;
                    const double vr = rowok ? vs[rr] : 0.0;
                    double a[kTW];
                    double acc = 0.0;
                    if (cb + kTW - 1 < rb && rb + 63 < n) {
                        // interior tile: strictly below the diagonal and inside the matrix,
                        // so every load is a uniform column base plus the lane's row
                        const unsigned ru = (unsigned)rr;
#pragma unroll
                        for (int qq = 0; qq < kTW; ++qq) a[qq] = (A + (size_t)(cb + qq) * ld)[ru];
#pragma unroll
                        for (int qq = 0; qq < kTW; ++qq) {
                            acc += a[qq] * vs[cb + qq];
                            a[qq] *= vr;
                        }
                    } else {
                        // diagonal or edge tile: clamped addresses, masked values
                        const unsigned rc = (unsigned)min(rr, n - 1);
#pragma unroll
                        for (int qq = 0; qq < kTW; ++qq) {
                            const int c = cb + qq;
                            const double val = (A + (size_t)min(c, n - 1) * ld)[rc];
                            a[qq] = (rowok && c < n && rr >= c) ? val : 0.0;
                        }
#pragma unroll
                        for (int qq = 0; qq < kTW; ++qq) {
                            const int c = cb + qq;
                            acc += a[qq] * vs[min(c, n - 1)];
                            a[qq] = rr > c ? a[qq] * vr : 0.0;
                        }
                    }
                    const double tot = reduce16(a, lane);
                    const int c = cb + reduce16_index(lane);
                    if (lane < kTW && c < n) slice[c] += tot;
                    if (rowok) slice[rr] += acc;
                    u += kTriWaves;
                }
            }
            if (j > 0) {
                double py[16];
#pragma unroll
                for (int i = 0; i < kNB; ++i) {
                    const bool on = i < j && r >= c0 && r < n;
                    py[i] = on ? wreg[i] * v : 0.0;        // W_i^T v
                    py[kNB + i] = on ? vreg[i] * v : 0.0;  // V_i^T v
                }
                const double tot = reduce16(py, lane);
                if (lane < 16) ysp[wid * 16 + reduce16_index(lane)] = tot;
            }
            __syncthreads();
            PT_MARK(1);
            // ---- p = tau (A v - V W^T v - W V^T v); w = p - tau/2 (p.v) v
            double p = 0.0;
            if (r >= c0 && r < n) {
#pragma unroll
                for (int w = 0; w < kTriWaves; ++w) p += part[w * ld + r];
#pragma unroll
                for (int i = 0; i < kNB; ++i) {
                    if (i < j) {
                        double y1 = 0.0, y2 = 0.0;
#pragma unroll
                        for (int w = 0; w < kTriWaves; ++w) {
                            y1 += ysp[w * 16 + i];
                            y2 += ysp[w * 16 + kNB + i];
                        }
                        p -= vreg[i] * y1 + wreg[i] * y2;
                    }
                }
                p *= tk;
            }
            const double a2 = -0.5 * tk * bsum(p * v, red);
            const double w = p + a2 * v;
#pragma unroll
            for (int i = 0; i < kNB; ++i)
                if (i == j) wreg[i] = w;
            if (r == k + 1 && j + 1 < b) {
#pragma unroll
                for (int i = 0; i < kNB; ++i) {
                    rowk[i] = vreg[i];
                    rowk[kNB + i] = wreg[i];
                }
            }
            __syncthreads();
            PT_MARK(2);
        }
        // ---- rank-2b update of the trailing lower triangle (columns >= cs)
        const int cs = k0 + b;
        if (cs < n) {
            if (r < n) {
#pragma unroll
                for (int i = 0; i < kNB; ++i) {
                    Vst[(size_t)r * kVS + i] = vreg[i];
                    Wst[(size_t)r * kVS + i] = wreg[i];
                }
            } else if (r < ld) {
#pragma unroll
                for (int i = 0; i < kNB; ++i) Vst[(size_t)r * kVS + i] = Wst[(size_t)r * kVS + i] = 0.0;
            }
            __syncthreads();
            // A -= V W^T + W V^T on the FP64 matrix cores, in 16 x 16 tiles on or
            // below the diagonal from row / column T0 = cs rounded down to 16.  A
            // tile sits in the accumulator transposed, element e of lane (li, lg)
            // being A(R0 + li, C0 + 4 e + lg), so its loads and stores are
            // column-contiguous; the update is D += Aop Bop with
            // Aop = -[W_C | V_C] (16 x 2 kNB) and Bop = [V_R ; W_R] (2 kNB x 16).
            // A wavefront takes strips of up to four tiles in one column block
            // (sharing Aop) and issues their 16 loads together.
            static_assert(kNB == 8, "two MFMA k-steps per panel factor");
            const int T0 = cs & ~15;
            const int nt16 = (n - T0 + 15) >> 4;
            const int li = lane & 15, lg = lane >> 4;
            for (int q = wid;; q += kTriWaves) {
                int tc = 0, rem = q;
                while (tc < nt16) {
                    const int ns = (nt16 - tc + 3) >> 2;
                    if (rem < ns) break;
                    rem -= ns;
                    ++tc;
                }
                if (tc >= nt16) break;
                const int C0 = T0 + 16 * tc, Rb = C0 + 64 * rem;
                const int ntl = min(4, nt16 - tc - 4 * rem);
                double aop[4];
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2)
                    aop[s2] = -(s2 < 2 ? Wst[(size_t)(C0 + li) * kVS + 4 * s2 + lg]
                                       : Vst[(size_t)(C0 + li) * kVS + 4 * (s2 - 2) + lg]);
                tvec4 acc[4];
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int e2 = 0; e2 < 4; ++e2) {
                        const int row = Rb + 16 * t + li, col = C0 + 4 * e2 + lg;
                        const bool ok = t < ntl && row < n && col >= cs && row >= col;
                        acc[t][e2] = ok ? A[row + (size_t)col * ld] : 0.0;
                    }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (t < ntl) {
                        const int R0 = Rb + 16 * t;
#pragma unroll
                        for (int s2 = 0; s2 < 4; ++s2) {
                            const double bop = s2 < 2 ? Vst[(size_t)(R0 + li) * kVS + 4 * s2 + lg]
                                                      : Wst[(size_t)(R0 + li) * kVS + 4 * (s2 - 2) + lg];
                            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop[s2], bop, acc[t], 0, 0, 0);
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int e2 = 0; e2 < 4; ++e2) {
                        const int row = Rb + 16 * t + li, col = C0 + 4 * e2 + lg;
                        if (t < ntl && row < n && col >= cs && row >= col) A[row + (size_t)col * ld] = acc[t][e2];
                    }
            }
            __syncthreads();
        }
        PT_MARK(3);
    }
    PT_STORE(0);
    if (tid == 0 && n > 0) {
        d[n - 1] = A[(size_t)(n - 1) * ld + (n - 1)];
        e[n - 1] = 0.0;
        tau[n - 1] = 0.0;
    }
}

// Numbers of eigenvalues of T below x0 and below x1 (Sturm counts) from the
// signs of the leading principal minors of T - x I,
//     p_i = (d_i - x) p_{i-1} - e_{i-1}^2 p_{i-2},   p_{-1} = 1,
// counting i where p_i and p_{i-1} differ in sign (the pivot q_i = p_i/p_{i-1}
// of the LDL^T recurrence is negative).  No divisions: one multiply-add per step
// on each chain, and the two shifts run as independent chains.  Both minors are
// rescaled by a power of two every eight steps (signs and ratios unchanged).  A
// minor that is exactly zero counts as non-negative (that only decides on which
// side of an exact eigenvalue of a leading block the shift is counted) and is
// replaced by a tiny negative pivot at the next rescaling.
__device__ __forceinline__ void sturm_count2(const double* d, const double* e2, int n, double x0, double x1,
                                             int& c0, int& c1) {
    double p0 = d[0] - x0, q0 = 1.0;  // p_i, p_{i-1}
    double p1 = d[0] - x1, q1 = 1.0;
    int k0 = p0 < 0.0, k1 = p1 < 0.0;
    auto step = [&](double di, double ei) {
        const double n0 = fma(di - x0, p0, -ei * q0);
        const double n1 = fma(di - x1, p1, -ei * q1);
        k0 += (n0 < 0.0) != (p0 < 0.0);
        k1 += (n1 < 0.0) != (p1 < 0.0);
        q0 = p0;
        p0 = n0;
        q1 = p1;
        p1 = n1;
    };
    int i = 1;
    // eight steps per block: the sixteen LDS reads are issued before the arithmetic
    for (; i + 8 <= n; i += 8) {
        double dd[8], ee[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            dd[u] = d[i + u];
            ee[u] = e2[i + u - 1];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) step(dd[u], ee[u]);
        // an exact zero minor followed by a zero off-diagonal would stall the
        // recurrence: treat it as a tiny negative pivot (LAPACK's -pivmin)
        if (p0 == 0.0) p0 = (q0 == 0.0 ? -1.0 : -q0) * 1e-280;
        if (p1 == 0.0) p1 = (q1 == 0.0 ? -1.0 : -q1) * 1e-280;
        const int s0 = -ilogb(p0), s1 = -ilogb(p1);
        p0 = ldexp(p0, s0);
        q0 = ldexp(q0, s0);
        p1 = ldexp(p1, s1);
        q1 = ldexp(q1, s1);
    }
    for (; i < n; ++i) step(d[i], e2[i - 1]);
    c0 = k0;
    c1 = k1;
}

// deterministic pseudo-random start vector entry in (-1, 1)
__device__ __forceinline__ double start_value(int i, int j) {
    unsigned int h = (unsigned int)(i * 2654435761u) ^ (unsigned int)((j + 1) * 40503u) ^ 0x9e3779b9u;
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    h ^= h >> 12;
    h *= 0x297a2d39u;
    h ^= h >> 15;
    return (double)(h & 0xffffff) / 8388608.0 - 1.0;
}

// Cyclic Jacobi for a small symmetric matrix H (m x m, row stride K, single
// thread).  On exit R holds the eigenvectors as columns, ordered by ascending
// eigenvalue.
__device__ void jacobi_small(double* H, double* R, int m, int K) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) R[i * K + j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                const double h2 = H[i * K + j] * H[i * K + j];
                tot += h2;
                if (i != j) off += h2;
            }
        if (off <= 1e-34 * tot) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = H[p * K + q];
                if (apq == 0.0) continue;
                const double theta = (H[q * K + q] - H[p * K + p]) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < m; ++k) {
                    const double hp = H[k * K + p], hq = H[k * K + q];
                    H[k * K + p] = c * hp - s * hq;
                    H[k * K + q] = s * hp + c * hq;
                }
                for (int k = 0; k < m; ++k) {
                    const double hp = H[p * K + k], hq = H[q * K + k];
                    H[p * K + k] = c * hp - s * hq;
                    H[q * K + k] = s * hp + c * hq;
                }
                for (int k = 0; k < m; ++k) {
                    const double rp = R[k * K + p], rq = R[k * K + q];
                    R[k * K + p] = c * rp - s * rq;
                    R[k * K + q] = s * rp + c * rq;
                }
            }
    }
    for (int i = 0; i < m - 1; ++i) {
        int best = i;
        for (int j = i + 1; j < m; ++j)
            if (H[j * K + j] < H[best * K + best]) best = j;
        if (best != i) {
            const double hb = H[best * K + best];
            H[best * K + best] = H[i * K + i];
            H[i * K + i] = hb;
            for (int k = 0; k < m; ++k) {
                const double r = R[k * K + i];
                R[k * K + i] = R[k * K + best];
                R[k * K + best] = r;
            }
        }
    }
}

// Eigenvector of T for an isolated eigenvalue x by a twisted factorisation
// (K. V. Fernando, SIAM J. Matrix Anal. Appl. 18 (1997) 1013; Dhillon & Parlett,
// Linear Algebra Appl. 387 (2004) 1):
//     T - x I = L+ D+ L+^T (top down) = U- D- U-^T (bottom up),
// gamma_r = D+_r - b_r U-_r; at the twist index r of smallest |gamma_r| the
// vector solving N_r Delta_r N_r^T z = gamma_r e_r is z_r = 1,
// z_i = -L+_i z_{i+1} (i < r), z_i = -U-_{i-1} z_{i-1} (i > r), L+_i = b_i / D+_i,
// U-_i = b_i / D-_{i+1}.  Three passes over per-vector columns of Dp, Lp, Um
// and Z (element i at i * kcap); the up-solve runs inside the bottom-up pass,
// restarted whenever a better twist index appears.  Returns false if the
// vector overflowed (the caller then uses inverse iteration).
__device__ bool twisted_vector(const double* D, const double* Eo, int n, double x, double pivmin, int kcap,
                               gptr<double> Z, double* __restrict__ Dp, double* __restrict__ Lp,
                               double* __restrict__ Um) {
    double dp = D[0] - x;
    for (int i = 0; i < n - 1; ++i) {
        if (fabs(dp) < pivmin) dp = -pivmin;
        const size_t o = (size_t)i * kcap;
        const double lf = Eo[i] / dp;
        Dp[o] = dp;
        Lp[o] = lf;
        dp = (D[i + 1] - x) - lf * Eo[i];
    }
    if (fabs(dp) < pivmin) dp = -pivmin;
    double dm = D[n - 1] - x, gbest = dp, zr = 1.0, up2 = 1.0;
    int r = n - 1;
    Z[(size_t)(n - 1) * kcap] = 1.0;
    // the passes below read back what the first one stored: kLB loads are issued
    // together ahead of the dependent steps that use them
    for (int i0 = n - 2; i0 >= 0; i0 -= kLB) {
        double dpv[kLB], lpv[kLB];
#pragma unroll
        for (int u = 0; u < kLB; ++u) {
            const size_t o = (size_t)max(i0 - u, 0) * kcap;
            dpv[u] = Dp[o];
            lpv[u] = Lp[o];
        }
#pragma unroll
        for (int u = 0; u < kLB; ++u) {
            const int i = i0 - u;
            if (i < 0) break;
            if (fabs(dm) < pivmin) dm = -pivmin;
            const size_t o = (size_t)i * kcap;
            const double uu = Eo[i] / dm;
            Um[o] = uu;
            const double g = dpv[u] - Eo[i] * uu;
            dm = (D[i] - x) - uu * Eo[i];
            double zi;
            if (fabs(g) < fabs(gbest)) {
                gbest = g;
                r = i;
                zi = 1.0;
                up2 = 0.0;
            } else {
                zi = -lpv[u] * zr;
            }
            Z[o] = zi;
            up2 += zi * zi;
            zr = zi;
        }
    }
    double zd = 1.0, dn2 = 0.0;
    for (int i0 = r + 1; i0 < n; i0 += kLB) {
        double umv[kLB];
#pragma unroll
        for (int u = 0; u < kLB; ++u) umv[u] = Um[(size_t)(min(i0 + u, n - 1) - 1) * kcap];
#pragma unroll
        for (int u = 0; u < kLB; ++u) {
            const int i = i0 + u;
            if (i >= n) break;
            zd = -umv[u] * zd;
            Z[(size_t)i * kcap] = zd;
            dn2 += zd * zd;
        }
    }
    const double nrm2 = up2 + dn2;
    if (!(nrm2 > 0.0 && nrm2 < INFINITY)) return false;
    const double inv = 1.0 / sqrt(nrm2);
    for (int i0 = 0; i0 < n; i0 += kLB) {
        double zv[kLB];
#pragma unroll
        for (int u = 0; u < kLB; ++u) zv[u] = Z[(size_t)min(i0 + u, n - 1) * kcap];
#pragma unroll
        for (int u = 0; u < kLB; ++u)
            if (i0 + u < n) Z[(size_t)(i0 + u) * kcap] = zv[u] * inv;
    }
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// Eigenvalues, occupations (electronic.rs:3853-3921) and occupied eigenvectors
// of T.  LDS: 5 ld + 2 kEigMaxCluster^2 + 16 doubles + 3 ld ints.
__global__ void __launch_bounds__(kThreads) k_tridiag_eig(const int* act, const int* nvec, int ld, const double* dvec,
                                                          const double* evec, double* eps_all, const int64_t* vec0,
                                                          double* occ_all, SccState* st, double* const* Ztptr,
                                                          double* lu, int kcap, int full_spectrum) {
    extern __shared__ double sm[];
    const int l = blockIdx.x;
    const int inst = act[l];
    if (inst < 0) return;
    const int n = nvec[inst];
    const double* dg = dvec + (size_t)l * ld;
    const double* eg = evec + (size_t)l * ld;
    double* D = sm;             // diagonal
    double* E2 = sm + ld;       // squared off-diagonals
    double* Eo = sm + 2 * ld;   // off-diagonals
    double* lam = sm + 3 * ld;  // eigenvalues, ascending
    double* sh = sm + 4 * ld;   // inverse-iteration shifts
    constexpr int KC = kEigMaxCluster;
    double* Hc = sm + 5 * ld;   // projected cluster matrix
    double* Rc = Hc + KC * KC;  // its eigenvectors
    double* red = Rc + KC * KC;
    int* tight = (int*)(red + 16);  // 1 where lam[j] joins the group of lam[j-1]
    int* vflag = tight + ld;         // 1 where eigenvector j is formed by inverse iteration
    int* order = vflag + ld;         // eigenvector indices of the current pass
    const int tid = threadIdx.x, nt = blockDim.x;
    PT_INIT();
    for (int i = tid; i < n; i += nt) {
        D[i] = dg[i];
        Eo[i] = i < n - 1 ? eg[i] : 0.0;
        E2[i] = i < n - 1 ? eg[i] * eg[i] : 0.0;
    }
    __syncthreads();
    // Gershgorin interval and pivot guard (LAPACK dstebz conventions)
    double lo = INFINITY, hi = -INFINITY, emax = 0.0, tnorm = 0.0;
    for (int i = tid; i < n; i += nt) {
        const double r = (i > 0 ? fabs(Eo[i - 1]) : 0.0) + (i < n - 1 ? fabs(Eo[i]) : 0.0);
        lo = fmin(lo, D[i] - r);
        hi = fmax(hi, D[i] + r);
        emax = fmax(emax, i < n - 1 ? E2[i] : 0.0);
        tnorm = fmax(tnorm, fabs(D[i]) + r);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_down(lo, off, 64));
        hi = fmax(hi, __shfl_down(hi, off, 64));
        emax = fmax(emax, __shfl_down(emax, off, 64));
        tnorm = fmax(tnorm, __shfl_down(tnorm, off, 64));
    }
    __shared__ double s4[4][16];
    const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
    if (lane == 0) {
        s4[0][wid] = lo;
        s4[1][wid] = hi;
        s4[2][wid] = emax;
        s4[3][wid] = tnorm;
    }
    __syncthreads();
    lo = s4[0][0];
    hi = s4[1][0];
    emax = s4[2][0];
    tnorm = s4[3][0];
    for (int q = 1; q < nw; ++q) {
        lo = fmin(lo, s4[0][q]);
        hi = fmax(hi, s4[1][q]);
        emax = fmax(emax, s4[2][q]);
        tnorm = fmax(tnorm, s4[3][q]);
    }
    __syncthreads();
    const double pivmin = fmax(kSafeMin * fmax(1.0, emax), 1e-300);
    const double widen = 2.0 * kEps * fmax(fabs(lo), fabs(hi)) + 4.0 * pivmin;
    lo -= widen;
    hi += widen;
    // ---- eigenvalues by Sturm-count trisection, one per thread (the two count
    //      chains split the bracket into thirds).  The result arrays of a single
    //      system need every orbital energy; otherwise only the lowest m are
    //      formed (the occupied ones plus a margin) and the set is extended to all
    //      n when the Fermi window reaches the last computed eigenvalue.
    SccState& s = st[inst];
    const double kt = s.kt, nelec = s.nelec;
    double* eps = eps_all + vec0[inst];
    double* occ = occ_all + vec0[inst];
    int m = full_spectrum ? n : min(n, (int)ceil(0.5 * nelec) + kEigMargin);
    int have = 0;
    double mu = 0.0, ts = 0.0;
    int kneed = 0;
    // occupation and entropy term of an orbital 80 kT or more above the Fermi level:
    // the exponent is clamped there, so every such orbital contributes exactly these
    const double f_tail = 2.0 / (1.0 + exp(80.0));
    double ts_tail = 0.0;
    {
        const double h = fmin(1.0 - 1e-16, fmax(1e-16, 0.5 * f_tail));
        ts_tail = 2.0 * kt * (h * log(h) + (1.0 - h) * log(1.0 - h));
    }
    // From the second SCC iteration on, eps holds this instance's eigenvalues of
    // the previous iteration (the ones not formed then hold the Gershgorin bound):
    // a bracket of +-1e-4 |T| around the previous value is used when two Sturm
    // counts confirm that it holds eigenvalue j, which saves about a quarter of
    // the trisection steps; the converged value meets the same tolerance.
    const bool warm = s.total_iter > 0;
    const double dl = 1e-4 * tnorm + pivmin;
    for (;;) {
        for (int j = have + tid; j < m; j += nt) {
            double a = lo, b = hi;
            if (warm) {
                const double x0 = eps[j];
                if (x0 - dl > lo && x0 + dl < hi) {
                    int c0, c1;
                    sturm_count2(D, E2, n, x0 - dl, x0 + dl, c0, c1);
                    if (c0 <= j && c1 > j) {
                        a = x0 - dl;
                        b = x0 + dl;
                    }
                }
            }
            for (int it = 0; it < 200; ++it) {
                const double w = b - a;
                const double x0 = a + w * (1.0 / 3.0), x1 = a + w * (2.0 / 3.0);
                if (!(x0 > a && x1 > x0 && b > x1) || w <= 2.0 * kEps * fmax(fabs(a), fabs(b)) + pivmin) break;
                int c0, c1;
                sturm_count2(D, E2, n, x0, x1, c0, c1);
                if (c0 > j) b = x0;
                else if (c1 > j) {
                    a = x0;
                    b = x1;
                } else a = x1;
            }
            lam[j] = 0.5 * (a + b);
        }
        __syncthreads();
        PT_MARK(0);
        const int tail = n - m;  // orbitals not formed, all >= 80 kT above mu
        // ---- Fermi occupations (same semantics as k_fermi)
        const double mn = lam[0], mx = tail ? hi : lam[n - 1];
        if (kt <= 0.0) {
            if (tid == 0) {
                double rem = fmax(nelec, 0.0);
                int homo = 0;
                for (int i = 0; i < ld; ++i) {
                    double f = 0.0;
                    if (i < n && rem > 0.0) {
                        f = fmin(rem, 2.0);
                        rem -= f;
                    }
                    occ[i] = f;
                    if (i < n && f > 1e-12) homo = i;
                }
                mu = lam[homo];
                kneed = homo + 1;
                red[0] = mu;
                red[1] = (double)kneed;
            }
            __syncthreads();
            mu = red[0];
            kneed = (int)red[1];
            __syncthreads();
            break;
        } else if (nelec <= 0.0 || nelec >= 2.0 * n) {
            const double f = nelec <= 0.0 ? 0.0 : 2.0;
            for (int i = tid; i < ld; i += nt) occ[i] = i < n ? f : 0.0;
            mu = nelec <= 0.0 ? lam[0] : lam[n - 1];
            kneed = nelec <= 0.0 ? 0 : n;
            break;
        }
        double a = mn - 100.0 * kt - 10.0, b = mx + 100.0 * kt + 10.0;
        for (int it = 0; it < 200; ++it) {
            const double mid = 0.5 * (a + b);
            double part = 0.0;
            for (int i = tid; i < m; i += nt) {
                const double xx = fmin(80.0, fmax(-80.0, (lam[i] - mid) / kt));
                part += 2.0 / (1.0 + exp(xx));
            }
            if (tid == 0) part += tail * f_tail;
            const double sum = bsum(part, red);
            double na = a, nb = b;
            if (sum < nelec) na = mid;
            else nb = mid;
            if (na == a && nb == b) break;
            a = na;
            b = nb;
        }
        mu = 0.5 * (a + b);
        PT_MARK(1);
        if (tail && lam[m - 1] - mu < 80.0 * kt) {
            // the smearing reaches beyond the formed eigenvalues: form the rest
            have = m;
            m = n;
            continue;
        }
        double tpart = tid == 0 ? tail * ts_tail : 0.0;
        int kpart = 0;
        for (int i = tid; i < ld; i += nt) {
            double f = 0.0;
            if (i < m) {
                const double xx = fmin(80.0, fmax(-80.0, (lam[i] - mu) / kt));
                f = 2.0 / (1.0 + exp(xx));
                const double h = fmin(1.0 - 1e-16, fmax(1e-16, 0.5 * f));
                tpart += 2.0 * kt * (h * log(h) + (1.0 - h) * log(1.0 - h));
                if (f >= 1e-28) kpart = max(kpart, i + 1);
            } else if (i < n) {
                f = f_tail;
            }
            occ[i] = f;
        }
        ts = bsum(tpart, red);
        int km = kpart;
        for (int off = 32; off > 0; off >>= 1) km = max(km, __shfl_down(km, off, 64));
        __shared__ int skm[16];
        if (lane == 0) skm[wid] = km;
        __syncthreads();
        km = skm[0];
        for (int q = 1; q < nw; ++q) km = max(km, skm[q]);
        kneed = km;
        __syncthreads();
        break;
    }
    // orbital energies that were not formed are reported as the Gershgorin bound
    // (they only reach the batch's internal arrays, with zero eigenvector columns)
    for (int i = tid; i < ld; i += nt) eps[i] = i < m ? lam[i] : (i < n ? hi : kPadEnergy);
    if (tid == 0) {
        s.mu = mu;
        s.ts = ts;
        s.kneed_full = kneed;
        s.kneed = min(kneed, kcap);
    }
    kneed = min(kneed, kcap);
    PT_MARK(2);
    // ---- groups of nearly equal eigenvalues and inverse-iteration shifts
    // (coincident shifts are separated by pertol as in LAPACK dstein)
    const double tn = fmax(tnorm, 1e-300);
    const double ctol = 1e-5 * tn, pertol = 10.0 * kEps * tn, tiny = kEps * tn;
    for (int j = tid; j < kneed; j += nt) tight[j] = (j > 0 && lam[j] - lam[j - 1] < ctol) ? 1 : 0;
    __syncthreads();
    if (tid == 0) {
        double prev = -INFINITY;
        for (int j = 0; j < kneed; ++j) {
            double x = lam[j];
            if (tight[j] && x - prev < pertol) x = prev + pertol;
            sh[j] = x;
            prev = x;
        }
    }
    __syncthreads();
    gptr<double> Z = (gptr<double>)Ztptr[l];
    const size_t fld = (size_t)ld * kcap;
    double* __restrict__ rU1 = lu + (size_t)l * 5 * fld;  // 1 / diagonal of U
    double* __restrict__ U2 = rU1 + fld;                  // first superdiagonal of U
    double* __restrict__ U3 = rU1 + 2 * fld;              // second superdiagonal (row interchanges)
    double* __restrict__ Lm = rU1 + 3 * fld;              // multipliers
    double* __restrict__ Pv = rU1 + 4 * fld;              // 1 where rows i, i+1 were interchanged
    // ---- eigenvectors, one thread each; element (i, j) of every per-vector array
    //      sits at i * kcap + j so a wavefront's accesses coalesce.  Isolated
    //      eigenvalues: twisted factorisation; members of groups (and any twisted
    //      vector that overflowed): inverse iteration from distinct start vectors,
    //      then the group treatment below.
    __shared__ int npass;
    for (int j = tid; j < kneed; j += nt) vflag[j] = (tight[j] || (j + 1 < kneed && tight[j + 1])) ? 1 : 0;
    __syncthreads();
    if (tid == 0) {
        int c = 0;
        for (int j = 0; j < kneed; ++j)
            if (!vflag[j]) order[c++] = j;
        npass = c;
    }
    __syncthreads();
    for (int t = tid; t < npass; t += nt) {
        const int j = order[t];
        if (!twisted_vector(D, Eo, n, lam[j], pivmin, kcap, Z + j, rU1 + j, U2 + j, U3 + j)) vflag[j] = 1;
    }
    __syncthreads();
    PT_MARK(3);
    if (tid == 0) {
        int c = 0;
        for (int j = 0; j < kneed; ++j)
            if (vflag[j]) order[c++] = j;
        npass = c;
    }
    __syncthreads();
    for (int t = tid; t < npass; t += nt) {
        const int j = order[t];
        const double x = sh[j];
        double a = D[0] - x, bb = n > 1 ? Eo[0] : 0.0;
        for (int i = 0; i < n - 1; ++i) {
            const double sub = Eo[i], dn = D[i + 1] - x, en = i + 1 < n - 1 ? Eo[i + 1] : 0.0;
            const size_t o = (size_t)i * kcap + j;
            if (fabs(a) >= fabs(sub)) {
                const double u = fabs(a) < tiny ? copysign(tiny, a) : a;
                const double m = sub / u;
                rU1[o] = 1.0 / u;
                U2[o] = bb;
                U3[o] = 0.0;
                Lm[o] = m;
                Pv[o] = 0.0;
                a = dn - m * bb;
                bb = en;
            } else {
                const double m = a / sub;
                rU1[o] = 1.0 / sub;
                U2[o] = dn;
                U3[o] = en;
                Lm[o] = m;
                Pv[o] = 1.0;
                a = bb - m * dn;
                bb = -m * en;
            }
        }
        {
            const size_t o = (size_t)(n - 1) * kcap + j;
            const double u = fabs(a) < tiny ? copysign(tiny, a) : a;
            rU1[o] = 1.0 / u;
            U2[o] = 0.0;
            U3[o] = 0.0;
        }
        double scale = 1.0, ss = 1.0;
        for (int it = 0; it < 3; ++it) {
            // y = L^-1 P z  (row interchanges applied on the fly); the next kLB
            // entries of z and of the factors are loaded before the steps use them
            double cur = it == 0 ? start_value(0, j) : Z[j] * scale;
            for (int i0 = 0; i0 < n - 1; i0 += kLB) {
                double nxv[kLB], mv[kLB], pv[kLB];
#pragma unroll
                for (int u = 0; u < kLB; ++u) {
                    const int i = min(i0 + u, n - 2);
                    const size_t o = (size_t)i * kcap + j;
                    nxv[u] = it == 0 ? start_value(i + 1, j) : Z[o + kcap] * scale;
                    mv[u] = Lm[o];
                    pv[u] = Pv[o];
                }
#pragma unroll
                for (int u = 0; u < kLB; ++u) {
                    const int i = i0 + u;
                    if (i >= n - 1) break;
                    const double nx = nxv[u], m = mv[u];
                    double out;
                    if (pv[u] != 0.0) {
                        out = nx;
                        cur = cur - m * nx;
                    } else {
                        out = cur;
                        cur = nx - m * cur;
                    }
                    Z[(size_t)i * kcap + j] = out;
                }
            }
            Z[(size_t)(n - 1) * kcap + j] = cur;
            // z = U^-1 y
            double z1 = 0.0, z2 = 0.0, zmax = 0.0;
            ss = 0.0;
            for (int i0 = n - 1; i0 >= 0; i0 -= kLB) {
                double yv[kLB], u2v[kLB], u3v[kLB], rv[kLB];
#pragma unroll
                for (int u = 0; u < kLB; ++u) {
                    const size_t o = (size_t)max(i0 - u, 0) * kcap + j;
                    yv[u] = Z[o];
                    u2v[u] = U2[o];
                    u3v[u] = U3[o];
                    rv[u] = rU1[o];
                }
#pragma unroll
                for (int u = 0; u < kLB; ++u) {
                    const int i = i0 - u;
                    if (i < 0) break;
                    const double zi = (yv[u] - u2v[u] * z1 - u3v[u] * z2) * rv[u];
                    Z[(size_t)i * kcap + j] = zi;
                    z2 = z1;
                    z1 = zi;
                    zmax = fmax(zmax, fabs(zi));
                    ss += zi * zi;
                }
            }
            scale = 1.0 / zmax;
        }
        const double inv = 1.0 / sqrt(ss);
        for (int i = 0; i < n; ++i) Z[(size_t)i * kcap + j] *= inv;
    }
    __syncthreads();
    PT_MARK(4);
    // zero rows n..ld-1 and columns kneed..kcap-1
    for (size_t q = (size_t)n * kcap + tid; q < fld; q += nt) Z[q] = 0.0;
    const int zc = kcap - kneed;
    if (zc > 0)
        for (int q = tid; q < n * zc; q += nt) Z[(size_t)(q / zc) * kcap + kneed + q % zc] = 0.0;
    __syncthreads();
    // ---- groups: orthonormalise (modified Gram-Schmidt, twice) and rotate to the
    //      eigenvectors of T projected on the group (Rayleigh-Ritz)
    int g0 = 0;
    while (g0 < kneed) {
        int g1 = g0 + 1;
        while (g1 < kneed && tight[g1]) ++g1;
        const int cs = g1 - g0;
        if (cs > KC) {
            if (tid == 0) s.eig_fail = 1;
        } else if (cs >= 2) {
            for (int a = 0; a < cs; ++a) {
                gptr<double> za = Z + g0 + a;
                for (int pass = 0; pass < 2; ++pass)
                    for (int b2 = 0; b2 < a; ++b2) {
                        gptr<const double> zb = Z + g0 + b2;
                        double p = 0.0;
                        for (int r = tid; r < n; r += nt) p += zb[(size_t)r * kcap] * za[(size_t)r * kcap];
                        p = bsum(p, red);
                        for (int r = tid; r < n; r += nt) za[(size_t)r * kcap] -= p * zb[(size_t)r * kcap];
                    }
                double q = 0.0;
                for (int r = tid; r < n; r += nt) q += za[(size_t)r * kcap] * za[(size_t)r * kcap];
                const double inv = 1.0 / sqrt(bsum(q, red));
                for (int r = tid; r < n; r += nt) za[(size_t)r * kcap] *= inv;
            }
            __syncthreads();
            const int npair = cs * (cs + 1) / 2;
            for (int p = tid; p < npair; p += nt) {
                int a = 0, rem = p;
                while (rem >= cs - a) {
                    rem -= cs - a;
                    ++a;
                }
                const int b2 = a + rem;
                gptr<const double> za = Z + g0 + a;
                gptr<const double> zb = Z + g0 + b2;
                double h = 0.0;
                for (int r = 0; r < n; ++r) {
                    double tz = D[r] * zb[(size_t)r * kcap];
                    if (r > 0) tz += Eo[r - 1] * zb[(size_t)(r - 1) * kcap];
                    if (r < n - 1) tz += Eo[r] * zb[(size_t)(r + 1) * kcap];
                    h += za[(size_t)r * kcap] * tz;
                }
                Hc[a * KC + b2] = h;
                Hc[b2 * KC + a] = h;
            }
            __syncthreads();
            if (tid == 0) jacobi_small(Hc, Rc, cs, KC);
            __syncthreads();
            for (int r = tid; r < n; r += nt) {
                gptr<double> zr = Z + (size_t)r * kcap + g0;
                double* tr = rU1 + (size_t)r * kcap + g0;
                for (int b2 = 0; b2 < cs; ++b2) {
                    double v = 0.0;
                    for (int a = 0; a < cs; ++a) v += zr[a] * Rc[a * KC + b2];
                    tr[b2] = v;
                }
                for (int b2 = 0; b2 < cs; ++b2) zr[b2] = tr[b2];
            }
            __syncthreads();
        }
        g0 = g1;
    }
    PT_MARK(5);
    PT_STORE(1);
}

// ---------------------------------------------------------------------------
// Z <- Q Z, Q = H_0 H_1 ... H_{n-2}, applied as blocks I - V T V^T of kBT = 16
// reflectors (LAPACK dlarft forward/columnwise) from the last block to the first,
// on the FP64 matrix cores.  One work-group per (matrix, 32 eigenvectors); the
// 16 x 16 tiles of its Z columns stay in MFMA accumulator layout (lane l, element
// e: row 4e + l/16, column l%16) in registers for the whole transformation, wave
// w owning row tiles w, w + 8, ...  With that layout, k-step e of W = V^T Z takes
// element e of a Z tile as its B operand, and the update Z -= V (T W)
// accumulates straight into the tiles.  The block's reflectors are staged in
// LDS; W and the Gram matrix V^T V (for T) are summed over the wavefronts in a
// fixed order.  The chunks of one matrix are mapped onto one XCD (dispatch order
// is round-robin over the eight XCDs), so its reflectors are shared in that L2.
typedef double dvec4 __attribute__((ext_vector_type(4)));
// Eigenvectors per work-group (COLS) and row tiles per wavefront (TILES) are
// template parameters: TILES = 2 covers matrices up to 256 rows with half the
// accumulator registers, and the column width trades the number of
// work-groups against how often each re-stages the reflectors.  Each column's
// arithmetic (its row tiles per wavefront, the fixed-order sum over the
// wavefronts) does not depend on either, so every variant gives the same bits
// and the choice among them is a matter of speed only (launch_backtransform).
constexpr int kBTWaves = 8;                                // 512 threads
constexpr int kXcds = 8;
constexpr int kVbS = kBT + 1;  // LDS row stride of the staged reflectors (odd: conflict-free row writes)

__device__ __forceinline__ dvec4 mfma16x16x4(double a, double b, dvec4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// T factor (upper triangular, 16 x 16) of every reflector block of one matrix:
// T[i][i] = tau_i, T[i][j] = -tau_j sum_{c=i}^{j-1} T[i][c] G[c][j] with the Gram
// matrix G = V^T V accumulated on the matrix cores.  One wavefront per (block,
// matrix); the factors go to work[(l * (ld / 16) + b) * 256], row-major.
__global__ void __launch_bounds__(64) k_block_t(const int* act, const int* nvec, int ld, double* const* Aptr,
                                                const double* tauvec, double* work) {
    __shared__ double Gs[kBT * kBT];
    const int b = blockIdx.x, l = blockIdx.y;
    if (act[l] < 0) return;
    const int n = nvec[act[l]];
    const int lo = b * kBT;
    if (n < 2 || lo > n - 2) return;
    const int bb = min(kBT, n - 1 - lo), base = (lo + 1) & ~15;
    gptr<const double> A = (gptr<const double>)Aptr[l];
    const int lane = threadIdx.x, li = lane & 15, lg = lane >> 4, kk = lo + li;
    dvec4 ag = {0.0, 0.0, 0.0, 0.0};
    for (int row0 = base; row0 < n; row0 += 16) {
        double va[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = row0 + 4 * e + lg;
            va[e] = (li < bb && r < n && r > kk) ? (r == kk + 1 ? 1.0 : A[r + (size_t)kk * ld]) : 0.0;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) ag = mfma16x16x4(va[e], va[e], ag);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) Gs[(4 * e + lg) * kBT + li] = ag[e];
    __syncthreads();
    if (lane < kBT) {
        const int i = lane;
        const double* tau = tauvec + (size_t)l * ld + lo;
        double tj[kBT], row[kBT];
#pragma unroll
        for (int j = 0; j < kBT; ++j) {
            tj[j] = j < bb ? tau[j] : 0.0;
            row[j] = 0.0;
        }
        if (i < bb) row[i] = tj[i];
#pragma unroll
        for (int j = 1; j < kBT; ++j) {
            if (j > i && j < bb) {
                double acc = 0.0;
#pragma unroll
                for (int c = 0; c < kBT; ++c)
                    if (c >= i && c < j) acc += row[c] * Gs[c * kBT + j];
                row[j] = -tj[j] * acc;
            }
        }
        double* out = work + ((size_t)l * (ld / 16) + b) * (kBT * kBT) + i * kBT;
#pragma unroll
        for (int j = 0; j < kBT; ++j) out[j] = row[j];
    }
}

template <int COLS, int TILES>
__global__ void __launch_bounds__(64 * kBTWaves) k_backtransform(const int* act, const int* nvec, int ld,
                                                                 double* const* Aptr, double* const* Ztptr,
                                                                 const SccState* st, int kcap, int count, int nch,
                                                                 const double* tfac) {
    constexpr int CT = COLS / 16;  // 16-column tiles per work-group
    extern __shared__ double sm[];
    double* Vb = sm;                                  // staged reflectors, (tile-aligned rows) x kVbS
    double* part = Vb + (size_t)ld * kVbS;            // T W partials (2 tiles) of half the wavefronts
    double* Wp = part + (kBTWaves / 2) * CT * 4 * 64;  // -(T W), 16 x COLS
    double* Tl = Wp + kBT * COLS;                  // T of the block, 16 x 16
    const int id = blockIdx.x, per = gridDim.x / kXcds;
    const int logical = (id % kXcds) * per + id / kXcds;
    const int l = logical / nch, chunk = logical % nch;
    if (l >= count) return;
    const int inst = act[l];
    if (inst < 0) return;
    const int n = nvec[inst];
    const int k = st[inst].kneed;
    const int c0 = chunk * COLS;
    if (c0 >= k || n < 2) return;
    gptr<const double> A = (gptr<const double>)Aptr[l];
    gptr<double> Z = (gptr<double>)Ztptr[l];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, li = lane & 15, lg = lane >> 4;
    dvec4 z[TILES][CT];
#pragma unroll
    for (int t = 0; t < TILES; ++t)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const int col = c0 + 16 * ct + li;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = 16 * (wid + kBTWaves * t) + 4 * e + lg;
                z[t][ct][e] = (r < n && col < kcap) ? Z[(size_t)r * kcap + col] : 0.0;
            }
        }
    const int tile_end = (n + 15) & ~15;
    for (int lo = ((n - 2) / kBT) * kBT; lo >= 0; lo -= kBT) {
        const int bb = min(kBT, n - 1 - lo), base = (lo + 1) & ~15;
        __syncthreads();
        // reflectors lo .. lo + bb - 1 on rows base .. tile_end - 1 (zero above each
        // reflector's unit entry and below row n - 1): thread t stages row base + t,
        // its 16 loads in flight together and each wavefront reading 64 consecutive
        // rows of a column; the block's T factor alongside
        if (tid < tile_end - base) {
            const int r = base + tid;
            double v[kBT];
#pragma unroll
            for (int i = 0; i < kBT; ++i) {
                const int kk = lo + i;
                v[i] = (i < bb && r < n && r > kk) ? (r == kk + 1 ? 1.0 : A[r + (size_t)kk * ld]) : 0.0;
            }
#pragma unroll
            for (int i = 0; i < kBT; ++i) Vb[tid * kVbS + i] = v[i];
        }
        if (tid < kBT * kBT) Tl[tid] = tfac[((size_t)l * (ld / 16) + lo / kBT) * (kBT * kBT) + tid];
        __syncthreads();
        // this wavefront's share of T W = T V^T Z: W over its row tiles (accumulator
        // layout, element e = reflectors 4e .. 4e + 3), then T W with that W as the B
        // operand of k-step e
        dvec4 aw[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) aw[ct] = dvec4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < TILES; ++t) {
            const int row0 = 16 * (wid + kBTWaves * t);
            if (row0 >= base && row0 < n) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const double va = Vb[(row0 + 4 * e + lg - base) * kVbS + li];
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct) aw[ct] = mfma16x16x4(va, z[t][ct][e], aw[ct]);
                }
            }
        }
        dvec4 tw[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) tw[ct] = dvec4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const double ta = Tl[li * kBT + 4 * e + lg];
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) tw[ct] = mfma16x16x4(ta, aw[ct][e], tw[ct]);
        }
        // partials summed in two steps through half-size LDS: the upper wavefronts
        // deposit theirs, the lower ones add their own in place
        const int slot = wid % (kBTWaves / 2);
        if (wid >= kBTWaves / 2) {
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int e = 0; e < 4; ++e) part[((slot * CT + ct) * 4 + e) * 64 + lane] = tw[ct][e];
        }
        __syncthreads();
        if (wid < kBTWaves / 2) {
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int e = 0; e < 4; ++e) part[((slot * CT + ct) * 4 + e) * 64 + lane] += tw[ct][e];
        }
        __syncthreads();
        // -(T W)[i][cc]: entry (i, j) of a tile sits in element i / 4 of lane j + 16 (i % 4)
        for (int q = tid; q < kBT * COLS; q += 64 * kBTWaves) {
            const int i = q / COLS, cc = q % COLS, ct = cc >> 4, j = cc & 15;
            const int off = (ct * 4 + (i >> 2)) * 64 + j + 16 * (i & 3);
            double sum = 0.0;
#pragma unroll
            for (int w = 0; w < kBTWaves / 2; ++w) sum += part[w * CT * 4 * 64 + off];
            Wp[i * COLS + cc] = -sum;
        }
        __syncthreads();
        // Z -= V (T W): k-step s covers reflectors 4s .. 4s + 3
        double wb[4][CT];
#pragma unroll
        for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) wb[s2][ct] = Wp[(4 * s2 + lg) * COLS + 16 * ct + li];
#pragma unroll
        for (int t = 0; t < TILES; ++t) {
            const int row0 = 16 * (wid + kBTWaves * t);
            if (row0 >= base && row0 < n) {
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2) {
                    const double va = Vb[(row0 + li - base) * kVbS + 4 * s2 + lg];
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct) z[t][ct] = mfma16x16x4(va, wb[s2][ct], z[t][ct]);
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < TILES; ++t)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const int col = c0 + 16 * ct + li;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = 16 * (wid + kBTWaves * t) + 4 * e + lg;
                if (r < n && col < kcap) Z[(size_t)r * kcap + col] = z[t][ct][e];
            }
        }
}

// ---------------------------------------------------------------------------
hipError_t launch_tridiag(hipStream_t s, int count, const int* act, const int* nvec, int ld, double* const* A,
                          double* d, double* e, double* tau) {
    if (ld > kEigMaxDim || ld > kTriThreads) return hipErrorInvalidValue;
    const size_t lds = ((size_t)2 * kVS * ld + kTriWaves * 16 + 16 + 2 * kNB) * sizeof(double);
    const int lds_max = (int)(((size_t)2 * kVS * kEigMaxDim + kTriWaves * 16 + 16 + 2 * kNB) * sizeof(double));
    static bool attr = false;  // LDS above 64 KiB must be requested explicitly
    if (!attr) {
        hipError_t err = hipFuncSetAttribute((const void*)k_tridiag, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             lds_max);
        if (err != hipSuccess) return err;
        attr = true;
    }
    hipLaunchKernelGGL(k_tridiag, dim3(count), dim3(kTriThreads), lds, s, act, nvec, ld, A, d, e, tau);
    return hipGetLastError();
}

hipError_t launch_tridiag_eig(hipStream_t s, int count, const int* act, const int* nvec, int ld, const double* d,
                              const double* e, double* eps_all, const int64_t* vec0, double* occ_all, SccState* st,
                              double* const* Zt, double* lu, int kcap, bool full_spectrum) {
    if (ld > kEigMaxDim || kcap > ld) return hipErrorInvalidValue;
    const size_t lds = ((size_t)5 * ld + 2 * kEigMaxCluster * kEigMaxCluster + 16) * sizeof(double) +
                       (size_t)3 * ld * sizeof(int);
    hipLaunchKernelGGL(k_tridiag_eig, dim3(count), dim3(kThreads), lds, s, act, nvec, ld, d, e, eps_all, vec0,
                       occ_all, st, Zt, lu, kcap, full_spectrum ? 1 : 0);
    return hipGetLastError();
}

namespace {
size_t backtransform_lds(int ld, int cols) {
    return ((size_t)ld * kVbS + (kBTWaves / 2) * (cols / 16) * 4 * 64 + kBT * cols + kBT * kBT) * sizeof(double);
}

template <int COLS, int TILES>
hipError_t launch_backtransform_t(hipStream_t s, int count, const int* act, const int* nvec, int ld,
                                  double* const* A, double* const* Zt, const SccState* st, int kcap, double* work) {
    static bool attr = false;  // LDS above 64 KiB must be requested explicitly
    if (!attr) {
        hipError_t err = hipFuncSetAttribute((const void*)k_backtransform<COLS, TILES>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)backtransform_lds(16 * kBTWaves * TILES, COLS));
        if (err != hipSuccess) return err;
        attr = true;
    }
    const int nch = (kcap + COLS - 1) / COLS;
    const int blocks = (count * nch + kXcds - 1) / kXcds * kXcds;
    hipLaunchKernelGGL((k_backtransform<COLS, TILES>), dim3(blocks), dim3(64 * kBTWaves), backtransform_lds(ld, COLS),
                       s, act, nvec, ld, A, Zt, st, kcap, count, nch, (const double*)work);
    return hipGetLastError();
}
}  // namespace

bool backtransform_width_ok(int ld, int cols) {
    return cols == 16 || cols == 32 || (cols == 64 && ld <= 16 * kBTWaves * 2);
}

hipError_t launch_backtransform(hipStream_t s, int count, const int* act, const int* nvec, int ld,
                                double* const* A, const double* tau, double* const* Zt, const SccState* st,
                                int kcap, double* work, int cols) {
    if (ld > kEigMaxDim || kcap > ld || ld % 16 != 0 || !backtransform_width_ok(ld, cols))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_block_t, dim3(ld / 16, count), dim3(64), 0, s, act, nvec, ld, A, tau, work);
    static_assert(kEigMaxDim == 16 * kBTWaves * 4, "four row tiles per wavefront cover the largest matrix");
    const bool small = ld <= 16 * kBTWaves * 2;
    switch (cols) {
        case 16:
            return small ? launch_backtransform_t<16, 2>(s, count, act, nvec, ld, A, Zt, st, kcap, work)
                         : launch_backtransform_t<16, 4>(s, count, act, nvec, ld, A, Zt, st, kcap, work);
        case 64:
            return launch_backtransform_t<64, 2>(s, count, act, nvec, ld, A, Zt, st, kcap, work);
        default:
            return small ? launch_backtransform_t<32, 2>(s, count, act, nvec, ld, A, Zt, st, kcap, work)
                         : launch_backtransform_t<32, 4>(s, count, act, nvec, ld, A, Zt, st, kcap, work);
    }
}

}  // namespace gfn1

#ifdef GFN1_PHASE_TIMING
// Profiling builds: copies and clears the phase ticks (kernel, matrix, phase).
extern "C" int gfn1_debug_phase_ticks(unsigned long long* out) {
    const size_t bytes = sizeof(gfn1::gfn1_phase_ticks);
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(gfn1::gfn1_phase_ticks), bytes, 0, hipMemcpyDeviceToHost) != hipSuccess)
        return 1;
    static unsigned long long zero[sizeof(gfn1::gfn1_phase_ticks) / sizeof(unsigned long long)] = {};
    return hipMemcpyToSymbol(HIP_SYMBOL(gfn1::gfn1_phase_ticks), zero, bytes, 0, hipMemcpyHostToDevice) != hipSuccess;
}
#endif
