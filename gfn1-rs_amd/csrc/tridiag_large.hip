// Householder tridiagonalisation of one large symmetric matrix over the whole
// device: LAPACK dsytrd with uplo = 'L' (dlatrd panels of kLargeTridiagPanel
// reflectors, the trailing matrix updated once per panel on the matrix cores).
//
// Column i of a panel (position j) takes two kernels.  With a_r the column
// after the panel's corrections, u_r = a_r for r >= i+2 (u_{i+1} = 0) and the
// reflector v = scale u + e_{i+1}, the symmetric product is linear in u:
// A v = scale (A u) + A(:, i+1), so it runs before the reflector is known.
//   k_lt_symv    y_u = A u over the trailing lower triangle (strips of 64 x 64
//                tiles, every partial in its own slot of `part`), the dot
//                products V^T u, W^T u and the norm |u|^2 per 256-row chunk;
//   k_lt_reduce  the reflector (dlarfg) from alpha = a_{i+1} and |u|^2, then per
//                row y = scale y_u + A(r, i+1), p = tau (y - V W^T v - W V^T v)
//                into W(:, j), v into A and V(:, j), partial sums of p.v, and
//                the next column after every panel correction but the one that
//                needs p.v.
// W(:, j) holds p until the next column's kernels add a2 v (a2 = -tau (p.v)/2)
// on the fly and write the finished values back (row i+1 excepted: only the
// trailing update reads it, and adds the term itself).  The symmetric product
// reads the trailing lower triangle once per column, across every CU.
#include "tridiag_large.hpp"

#include <algorithm>
#include <cmath>

#include "wave_reduce.hpp"

namespace gfn1 {
namespace {

constexpr int kNB = kLargeTridiagPanel;
static_assert(kNB == 32, "the dot products run as two groups of 16 panel columns");
constexpr int kTile = 64;
constexpr int kChunk = 256;        // rows per work-group of the dot products
constexpr int kRedRows = 64;       // rows per work-group of k_lt_reduce (four threads per row)
constexpr int kMaxPartials = 256;  // p.v partial sums per column: kLargeTridiagMaxDim / kRedRows
constexpr int kDot = 2 * kNB + 8;  // per chunk: V^T u (kNB), W^T u (kNB), |u|^2, padding
typedef double tvec4 __attribute__((ext_vector_type(4)));

// scalars of the current column, written by k_lt_symv for k_lt_reduce
enum { kScD = 0, kScAlpha, kScA2, kScS, kScalars };

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

// Sum of npv <= kMaxPartials partial sums: four per lane in order, then the
// wavefront; every wavefront forms the same sum in the same order.
__device__ __forceinline__ double partial_sum(const double* __restrict__ pv, int npv) {
    const int l = threadIdx.x & 63;
    double x = 0.0;
#pragma unroll
    for (int k = 0; k < kMaxPartials / 64; ++k)
        if (l + 64 * k < npv) x += pv[l + 64 * k];
    return wave_sum(x);
}

// a2 of the previous panel column (from its p.v partial sums) and
// s = p_i + 2 a2 v_i, the scalar of column i's last correction.  Every
// wavefront forms the same sums in the same order.
struct PrevColumn {
    double a2 = 0.0, s = 0.0;
};
__device__ __forceinline__ PrevColumn prev_column(int i, int j, const double* __restrict__ V,
                                                  const double* __restrict__ W, int ldv,
                                                  const double* __restrict__ pv, int npv,
                                                  const double* __restrict__ tau) {
    PrevColumn pc;
    pc.a2 = -0.5 * tau[i - 1] * partial_sum(pv, npv);
    pc.s = W[(size_t)(j - 1) * ldv + i] + 2.0 * pc.a2 * V[(size_t)(j - 1) * ldv + i];
    return pc;
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
constexpr int kMaxStrip = 8;
static_assert(kMaxStrip * 64 <= 2 * 256, "k_lt_symv stages a strip's u with two loads per thread");
int strip_length(int nt) {
    for (int strip = kMaxStrip; strip > 1; strip >>= 1)
        if (strips_before(nt, strip) >= 2048) return strip;
    return 1;
}

// y_u = A u over the trailing lower triangle.  A work-group takes one strip:
// lane l holds row R0 + l, wavefront w columns 16 w .. 16 w + 15 of each
// tile, and the next tile's 16 loads are in flight while the current one is
// used.  The row part (row r times the columns of the strip) is summed over
// the strip in registers; the column part of each tile (column c times the
// rows below the diagonal) by a transposed wavefront sum.  Slots: part[b][r]
// for b <= block(r) come from the row's strips (the strip total in its first
// block, zero in the others, plus the diagonal tile's column part in block(r)
// itself); part[b][r] for b > block(r) is the column part of tile (b,
// block(r)).  Every slot is written exactly once per column, so the sum over
// b is deterministic.  Work-groups from nstrips on form V^T u, W^T u and |u|^2
// over 256-row chunks; the first of them also records the column's scalars.
__global__ void __launch_bounds__(256)
    k_lt_symv(int n, int ld, int c0, int j, int fresh, const double* __restrict__ A, const double* __restrict__ V,
              const double* __restrict__ W, int ldv, const double* __restrict__ next,
              const double* __restrict__ pv, int npv, const double* __restrict__ tau, double* __restrict__ part,
              int ldp, double* __restrict__ dots, double* __restrict__ scal, double* __restrict__ a2s, int nstrips,
              int nt, int strip) {
    __shared__ double prs[4][kTile];
    __shared__ double pdc[kTile];
    __shared__ double uc[kMaxStrip * kTile];
    __shared__ double wdots[4][2 * kNB];
    __shared__ double wn[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int i = c0 - 1;
    const double* vp = fresh ? nullptr : V + (size_t)(j - 1) * ldv;
    const int t = blockIdx.x;
    // A strip's geometry, its first tile's loads and the loads behind u over
    // its rows and columns go out before the previous column's scalars, whose
    // sums they do not need.
    int RB = 0, CB0 = 0, CBe = 0, R0 = 0, r = 0;
    bool rok = false;
    double x[16];
    double un[3] = {0.0, 0.0, 0.0}, uv[3] = {0.0, 0.0, 0.0};  // u sources: columns tid, tid + 256; row r
    bool uok[3] = {false, false, false};
    auto load = [&](double (&xx)[16], int cb) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = cb * kTile + 16 * w + q;
            const bool ok = rok && c < n && c >= c0 && r >= c;
            xx[q] = ok ? A[r + (size_t)c * ld] : 0.0;
        }
    };
    if (t < nstrips) {
        int lo = 0, hi = nt - 1;  // the tile row of strip t: largest rho with strips_before(rho, strip) <= t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (strips_before(mid, strip) <= t) lo = mid;
            else hi = mid - 1;
        }
        const int b0 = c0 / kTile, rho = lo;
        RB = b0 + rho;
        CB0 = b0 + (t - strips_before(rho, strip)) * strip;
        CBe = min(CB0 + strip - 1, RB);
        R0 = RB * kTile;
        r = R0 + lane;
        rok = r < n && r >= c0;
        load(x, CB0);
        // the strip spans at most kMaxStrip * kTile = 2 * 256 columns
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) {
            const int k = tid + 256 * kk;
            const int c = kk < 2 ? CB0 * kTile + k : r;
            uok[kk] = (kk == 2 || k < (CBe - CB0 + 1) * kTile) && c > c0 && c < n;
            if (uok[kk]) {
                if (fresh) {
                    un[kk] = A[c + (size_t)i * ld];
                } else {
                    un[kk] = next[c];
                    uv[kk] = vp[c];
                }
            }
        }
    }
    PrevColumn pc;
    if (!fresh) pc = prev_column(i, j, V, W, ldv, pv, npv, tau);
    // column i after the panel's corrections (rows >= i)
    auto col = [&](int r) -> double { return fresh ? A[r + (size_t)i * ld] : next[r] - vp[r] * pc.s; };
    auto u_of = [&](int r) -> double { return (r > c0 && r < n) ? col(r) : 0.0; };
    // u from the loaded sources, as u_of forms it
    auto u_pre = [&](int kk) -> double { return uok[kk] ? (fresh ? un[kk] : un[kk] - uv[kk] * pc.s) : 0.0; };
    if (t >= nstrips) {
        const int ch = t - nstrips;
        if (ch == 0 && tid == 0) {
            scal[kScD] = col(i);
            scal[kScAlpha] = col(c0);
            scal[kScA2] = pc.a2;
            scal[kScS] = pc.s;
            if (!fresh) a2s[j - 1] = pc.a2;
        }
        const int r = c0 + ch * kChunk + tid;
        const double uu = u_of(r);
        // W(:, j-1) still holds p: a2 v is added on the fly
#pragma unroll
        for (int g = 0; g < 2 * kNB / 16; ++g) {
            const int jj0 = (g & 1) * 16;
            if (jj0 >= j) continue;  // uniform: depends on j only
            double cp[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int jj = jj0 + q;
                double m = 0.0;
                if (r < n && jj < j) {
                    m = g < 2 ? V[(size_t)jj * ldv + r] : W[(size_t)jj * ldv + r];
                    if (g >= 2 && jj == j - 1) m += pc.a2 * vp[r];
                }
                cp[q] = m * uu;
            }
            const double tot = wave::reduce16(cp, lane);
            if (lane < 16) wdots[w][g * 16 + wave::reduce16_index(lane)] = tot;
        }
        const double nn = wave_sum(uu * uu);
        if (lane == 0) wn[w] = nn;
        __syncthreads();
        if (tid < 2 * kNB) {
            const int jj = tid & (kNB - 1);
            if (jj < j)
                dots[(size_t)ch * kDot + tid] = ((wdots[0][tid] + wdots[1][tid]) + wdots[2][tid]) + wdots[3][tid];
        }
        if (tid == 0) dots[(size_t)ch * kDot + 2 * kNB] = ((wn[0] + wn[1]) + wn[2]) + wn[3];
        return;
    }
    const double ur = u_pre(2);
    // u over the strip's columns, once per work-group
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
        const int k = tid + 256 * kk;
        if (k < (CBe - CB0 + 1) * kTile) uc[k] = u_pre(kk);
    }
    __syncthreads();
    double xn[16];
    double rs = 0.0;
    for (int cb = CB0; cb <= CBe; ++cb) {
        if (cb < CBe) load(xn, cb + 1);
        const int C0 = cb * kTile + 16 * w;
        const double* ucb = uc + (cb - CB0) * kTile + 16 * w;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = C0 + q;
            rs += x[q] * ucb[q];
            x[q] = (cb != RB || r > c) ? x[q] * ur : 0.0;  // the column product, in place
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

__global__ void __launch_bounds__(256)
    k_lt_reduce(int n, int ld, int c0, int j, int fresh, int last_in_panel, int nblk, const double* __restrict__ part,
                int ldp, double* __restrict__ A, double* __restrict__ V, double* __restrict__ W, int ldv,
                double* __restrict__ next, const double* __restrict__ dots, int ndot,
                const double* __restrict__ scal, double* __restrict__ pv_out, double* __restrict__ d,
                double* __restrict__ e, double* __restrict__ tau) {
    __shared__ double dv[kNB], dw[kNB], wi[kNB], vi[kNB];
    __shared__ double s4[4][kDot];
    __shared__ double yq[4][kRedRows], cq[4][kRedRows], uq[4][kRedRows];
    __shared__ double red[4];
    const int tid = threadIdx.x;
    const int i = c0 - 1;
    const int q = tid >> 6;
    const int r = c0 + blockIdx.x * kRedRows + (tid & 63);
    const double* vp = fresh ? nullptr : V + (size_t)(j - 1) * ldv;
    // Loads that need nothing computed here are issued first, so that their
    // round trips overlap the chunk sums: row r's u source and column i+1 of
    // the trailing block (quarter 0), and row c0 of the panel's V and W.
    double u_a = 0.0, u_n = 0.0, u_v = 0.0, u_w = 0.0, acol = 0.0;
    if (q == 0 && r < n) {
        acol = A[r + (size_t)c0 * ld];
        if (fresh) {
            u_a = A[r + (size_t)i * ld];
        } else {
            u_n = next[r];
            u_v = vp[r];
            u_w = W[(size_t)(j - 1) * ldv + r];
        }
    }
    double w_c0 = 0.0, v_c0 = 0.0, vp_c0 = 0.0;
    if (tid < j) {
        w_c0 = W[(size_t)tid * ldv + c0];
        v_c0 = V[(size_t)tid * ldv + c0];
        if (!fresh && tid == j - 1) vp_c0 = vp[c0];
    }
    const double a2 = scal[kScA2], s = scal[kScS];
    {
        // chunk sums of V^T u, W^T u, |u|^2: four interleaved partial sums per
        // value, added in chunk order; loads eight chunks at a time (indices
        // clamped into the ndot >= 1 chunks, the extra values not added)
        const int idx = tid & 63;
        double acc = 0.0, accn = 0.0;
        const bool want = (idx & (kNB - 1)) < j;
        for (int ch0 = q; ch0 < ndot; ch0 += 32) {
            double la[8], ln[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const size_t row = (size_t)min(ch0 + 4 * u, ndot - 1) * kDot;
                la[u] = want ? dots[row + idx] : 0.0;
                ln[u] = idx == 0 ? dots[row + 2 * kNB] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (ch0 + 4 * u < ndot) {
                    acc += la[u];
                    accn += ln[u];
                }
        }
        s4[q][idx] = want ? acc : 0.0;
        if (idx == 0) s4[q][2 * kNB] = accn;
    }
    __syncthreads();
    // the reflector (dlarfg), the same in every work-group
    const double nrm2 = ((s4[0][2 * kNB] + s4[1][2 * kNB]) + s4[2][2 * kNB]) + s4[3][2 * kNB];
    const double dd = scal[kScD], alpha = scal[kScAlpha];
    double beta = alpha, tk = 0.0, scale = 0.0;
    if (nrm2 > 0.0) {
        const double nr = sqrt(alpha * alpha + nrm2);
        beta = alpha >= 0.0 ? -nr : nr;
        tk = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
    }
    // W(., jj) with the previous column's a2 v added, from its loaded values
    auto wf = [&](double w_raw, double vp_raw, int jj) -> double {
        return (!fresh && jj == j - 1) ? w_raw + a2 * vp_raw : w_raw;
    };
    if (tid < j) {
        const double vu = ((s4[0][tid] + s4[1][tid]) + s4[2][tid]) + s4[3][tid];
        const double wu = ((s4[0][kNB + tid] + s4[1][kNB + tid]) + s4[2][kNB + tid]) + s4[3][kNB + tid];
        wi[tid] = wf(w_c0, vp_c0, tid);
        vi[tid] = v_c0;
        dv[tid] = scale * vu + vi[tid];  // V^T v
        dw[tid] = scale * wu + wi[tid];  // W^T v
    }
    __syncthreads();
    if (blockIdx.x == 0 && tid == 0) {
        d[i] = dd;
        e[i] = beta;
        tau[i] = tk;
        A[i + (size_t)i * ld] = dd;
    }
    // four threads per row: quarter q adds the slots b = b0 + q, b0 + q + 4, ...
    // (alternately into two sums, eight loads in flight; slot indices clamped
    // below nblk, the extra values not added) and the panel columns
    // jj = q, q + 4, ... (all kNB / 4 loaded, those with jj >= j not added);
    // the quarters are combined in order
    if (r < n) {
        const int b0 = c0 / kTile;
        double y0 = 0.0, y1 = 0.0;
        for (int b = b0 + q; b < nblk; b += 32) {
            double ls[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) ls[u] = part[(size_t)min(b + 4 * u, nblk - 1) * ldp + r];
#pragma unroll
            for (int u = 0; u < 8; u += 2) {
                if (b + 4 * u < nblk) y0 += ls[u];
                if (b + 4 * u + 4 < nblk) y1 += ls[u + 1];
            }
        }
        double corr = 0.0, upd = 0.0;
        if (j > 0) {
            double lv[kNB / 4], lw[kNB / 4];
#pragma unroll
            for (int u = 0; u < kNB / 4; ++u) {
                lv[u] = V[(size_t)(q + 4 * u) * ldv + r];
                lw[u] = W[(size_t)(q + 4 * u) * ldv + r];
            }
            const double vpr = vp[r];
#pragma unroll
            for (int u = 0; u < kNB / 4; ++u) {
                const int jj = q + 4 * u;
                if (jj < j) {
                    const double vr = lv[u], wr = wf(lw[u], vpr, jj);
                    corr += vr * dw[jj] + wr * dv[jj];
                    upd += vr * wi[jj] + wr * vi[jj];
                }
            }
        }
        yq[q][tid & 63] = y0 + y1;
        cq[q][tid & 63] = corr;
        uq[q][tid & 63] = upd;
    }
    __syncthreads();
    double pvl = 0.0;
    if (q == 0 && r < n) {
        const int l = tid;
        const double yu = ((yq[0][l] + yq[1][l]) + yq[2][l]) + yq[3][l];
        const double corr = ((cq[0][l] + cq[1][l]) + cq[2][l]) + cq[3][l];
        const double upd = ((uq[0][l] + uq[1][l]) + uq[2][l]) + uq[3][l];
        const double u = r > c0 ? (fresh ? u_a : u_n - u_v * s) : 0.0;
        const double v = r == c0 ? 1.0 : scale * u;
        const double p = tk * (scale * yu + acol - corr);  // acol: column i+1 of the trailing block
        if (!fresh && r != c0) W[(size_t)(j - 1) * ldv + r] = wf(u_w, u_v, j - 1);  // finished
        W[(size_t)j * ldv + r] = p;
        V[(size_t)j * ldv + r] = v;
        A[r + (size_t)i * ld] = r == c0 ? beta : v;
        pvl = p * v;
        if (!last_in_panel) next[r] = acol - upd - p;  // v_{i+1} = 1
    }
    const double sum = group_sum<256>(pvl, red);
    if (tid == 0) pv_out[blockIdx.x] = sum;
}

// A(c:, c:) -= V W^T + W V^T on the lower triangle of the trailing block (m =
// n - c rows; A, V, W offset to row and column c), on the FP64 matrix cores,
// with the W terms still pending added on the fly: a2 v on the panel's last
// column (a2 from its p.v partials and tau_c = tau of that column) and, in row
// c, a2s[jb-2] v on the one before.  One work-group per 64 x 64 tile on or
// below the diagonal; wavefront w takes rows 16 w .. 16 w + 15 of the tile and
// its four 16 x 16 column blocks.  The block sits in the accumulator transposed
// (element e of lane (li, lg) is A(R0 + li, C0 + 4 e + lg)), so its loads and
// stores are column-contiguous; D' += Aop Bop over K = 2 kNB with
// Aop = -[W_C | V_C] and Bop = [V_R | W_R]^T.
__global__ void __launch_bounds__(256)
    k_lt_syr2k(int m, int jb, double* __restrict__ A, int ld, const double* __restrict__ V,
               const double* __restrict__ W, int ldv, const double* __restrict__ pv, int npv,
               const double* __restrict__ tau_c, const double* __restrict__ a2s) {
    const int t = blockIdx.x;
    int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= t) ++I;
    while (I * (I + 1) / 2 > t) --I;
    const int J = t - I * (I + 1) / 2;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lg = lane >> 4;
    const double a2_last = -0.5 * tau_c[0] * partial_sum(pv, npv);
    const double a2_prev = jb >= 2 ? a2s[jb - 2] : 0.0;
    auto wv = [&](int kc, int row) -> double {
        double x = W[(size_t)kc * ldv + row];
        if (kc == jb - 1) x += a2_last * V[(size_t)kc * ldv + row];
        else if (kc == jb - 2 && row == 0) x += a2_prev * V[(size_t)kc * ldv + row];
        return x;
    };
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
            const double a = (ok && cA < m) ? -(upper ? V[(size_t)kc * ldv + cA] : wv(kc, cA)) : 0.0;
            const double b = (ok && rB < m) ? (upper ? wv(kc, rB) : V[(size_t)kc * ldv + rB]) : 0.0;
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
    if ((err = hipMalloc(&dots, sizeof(double) * nch * kDot)) != hipSuccess) return err;
    if ((err = hipMalloc(&pv, sizeof(double) * 2 * (((size_t)rows + kRedRows - 1) / kRedRows))) != hipSuccess)
        return err;
    if ((err = hipMalloc(&next, sizeof(double) * rows)) != hipSuccess) return err;
    if ((err = hipMalloc(&scal, sizeof(double) * (kScalars + kNB))) != hipSuccess) return err;
    n = rows;
    return hipSuccess;
}

void LargeTridiagWork::release() {
    for (double** p : {&V, &W, &part, &dots, &pv, &next, &scal}) {
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
    const int nred_max = (w.n + kRedRows - 1) / kRedRows;
    double* scal = w.scal;
    double* a2s = w.scal + kScalars;
    int npv = 0;
    const double* pv_prev = w.pv;
    for (int i0 = 0; i0 < n - 1; i0 += kNB) {
        const int jb = std::min(kNB, n - 1 - i0);
        for (int j = 0; j < jb; ++j) {
            const int i = i0 + j, c0 = i + 1;
            // alternate p.v buffers: the previous column's partials stay readable
            double* pv_out = w.pv + (size_t)(i & 1) * nred_max;
            const int b0 = c0 / kTile, nt = nblk - b0;
            const int strip = strip_length(nt);
            const int nstrips = strips_before(nt, strip);
            const int nch = (n - c0 + kChunk - 1) / kChunk;
            const int fresh = j == 0 ? 1 : 0;
            hipLaunchKernelGGL(k_lt_symv, dim3(nstrips + nch), dim3(256), 0, s, n, ld, c0, j, fresh,
                               (const double*)A, (const double*)w.V, (const double*)w.W, ldv,
                               (const double*)w.next, pv_prev, npv, (const double*)tau, w.part, ldp, w.dots, scal,
                               a2s, nstrips, nt, strip);
            const int nred = (n - c0 + kRedRows - 1) / kRedRows;
            hipLaunchKernelGGL(k_lt_reduce, dim3(nred), dim3(256), 0, s, n, ld, c0, j, fresh,
                               j == jb - 1 ? 1 : 0, nblk, (const double*)w.part, ldp, A, w.V, w.W, ldv, w.next,
                               (const double*)w.dots, nch, (const double*)scal, pv_out, d, e, tau);
            pv_prev = pv_out;
            npv = nred;
        }
        // A(c:, c:) -= V W^T + W V^T
        const int c = i0 + jb;
        const int m = n - c, mt = (m + kTile - 1) / kTile;
        hipLaunchKernelGGL(k_lt_syr2k, dim3(mt * (mt + 1) / 2), dim3(256), 0, s, m, jb, A + c + (size_t)c * ld, ld,
                           (const double*)(w.V + c), (const double*)(w.W + c), ldv, pv_prev, npv,
                           (const double*)(tau + c - 1), (const double*)a2s);
    }
    hipLaunchKernelGGL(k_lt_last, dim3(1), dim3(1), 0, s, n, ld, (const double*)A, d, e, tau);
    return hipGetLastError();
}

}  // namespace gfn1
