// HIP kernels of the GFN1-xTB SCC energy + gradient path (gfx950).
//
// Layout: instances are stored instance-major in flat index tables (atoms,
// shells, AOs); dense matrices are per-instance column-major ld x ld blocks
// (ld = AO count padded to a multiple of 32).  Every kernel that walks the SCC
// takes a list of active instance ids so that converged instances drop out.
//
// Reference semantics are cited per kernel (paths relative to the reference
// repository root).
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "device.hpp"
#include "kernels.hpp"

namespace gfn1 {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrtPiK = 1.7724538509055160273;  // sqrt(pi)
constexpr double kPrimScreen = 1.0e-13;   // integrals.rs:17
constexpr double kPairEps2 = 1.0e-24;     // pairlist.rs:16 (DIST_EPS^2)

__device__ __forceinline__ double ipow(double x, int n) {
    double r = 1.0;
    for (int i = 0; i < n; ++i) r *= x;
    return r;
}

__device__ __forceinline__ double binom4(int n, int k) {
    // n <= 4
    const double t[5][5] = {{1, 0, 0, 0, 0}, {1, 1, 0, 0, 0}, {1, 2, 1, 0, 0}, {1, 3, 3, 1, 0}, {1, 4, 6, 4, 1}};
    return t[n][k];
}

// 1-D Gaussian-product overlap moment without the K_AB factor (binomial
// expansion about the product centre, integrals.rs:952-976).
__device__ double moment1d(int i, int j, double pa, double pb, const double* g) {
    double sum = 0.0;
    for (int u = 0; u <= i; ++u) {
        double ca = binom4(i, u) * ipow(pa, i - u);
        for (int v = 0; v <= j; ++v) {
            if ((u + v) & 1) continue;
            double cb = binom4(j, v) * ipow(pb, j - v);
            sum += ca * cb * g[(u + v) >> 1];
        }
    }
    return sum;
}

// moment1d in the arithmetic type T (FP32 on the mixed-precision gradient path).
template <typename T>
__device__ __forceinline__ T moment1d_t(int i, int j, T pa, T pb, const T* g) {
    T sum = 0;
    for (int u = 0; u <= i; ++u) {
        T ca = (T)binom4(i, u);
        for (int q = 0; q < i - u; ++q) ca *= pa;
        for (int v = 0; v <= j; ++v) {
            if ((u + v) & 1) continue;
            T cb = (T)binom4(j, v);
            for (int q = 0; q < j - v; ++q) cb *= pb;
            sum += ca * cb * g[(u + v) >> 1];
        }
    }
    return sum;
}

// Cartesian content of the real spherical AOs (basis.rs:382-441):
// s; px py pz; d x2-y2, z2, xy, xz, yz.
struct CartTerm {
    int8_t px, py, pz;
    double c;
};
__device__ __forceinline__ int ao_terms(int l, int ao, CartTerm* t) {
    const double h3 = 0.86602540378443864676;  // sqrt(3)/2
    const double s3 = 1.73205080756887729353;  // sqrt(3)
    if (l == 0) {
        t[0] = {0, 0, 0, 1.0};
        return 1;
    }
    if (l == 1) {
        t[0] = {(int8_t)(ao == 0), (int8_t)(ao == 1), (int8_t)(ao == 2), 1.0};
        return 1;
    }
    switch (ao) {
        case 0: t[0] = {2, 0, 0, h3}; t[1] = {0, 2, 0, -h3}; return 2;
        case 1: t[0] = {2, 0, 0, 0.5}; t[1] = {0, 2, 0, 0.5}; t[2] = {0, 0, 2, -1.0}; return 3;
        case 2: t[0] = {1, 1, 0, s3}; return 1;
        case 3: t[0] = {1, 0, 1, s3}; return 1;
        default: t[0] = {0, 1, 1, s3}; return 1;
    }
}
// sum of |component coefficients| of an AO (integrals.rs:21-24)
__device__ __forceinline__ double ao_abs_sum(int l, int ao) {
    if (l < 2) return 1.0;
    return ao == 1 ? 2.0 : 1.73205080756887729353;
}

__device__ __forceinline__ double kshell(const Globals& g, int l) { return l == 0 ? g.ks : l == 1 ? g.kp : g.kd; }
__device__ __forceinline__ double kpair(const Globals& g, int la, int lb) {
    if ((la == 0 && lb == 1) || (la == 1 && lb == 0)) return g.ksp;
    return 0.5 * (kshell(g, la) + kshell(g, lb));
}
// H0 shell-pair scaling (hamiltonian.rs:140-165)
__device__ double hscale(const DeviceTables& t, int za, const ShellType& A, int zb, const ShellType& B) {
    const Globals& g = t.glob;
    if (A.valence && B.valence) {
        const double den = t.elements[za].en - t.elements[zb].en;
        return t.pair_scale[za * (kMaxZ + 1) + zb] * kpair(g, A.l, B.l) * (1.0 + g.enscale * 0.01 * (den * den));
    }
    if (A.valence) return 0.5 * (kpair(g, A.l, A.l) + g.kdiff);
    if (B.valence) return 0.5 * (kpair(g, B.l, B.l) + g.kdiff);
    return g.kdiff;
}

// Exponential counting function and its radial derivative (coordination.rs:96-106).
__device__ __forceinline__ void count_fn(double r, double rc, double& value, double& dvalue) {
    const double raw = -16.0 * (rc / r - 1.0);
    const double arg = fmin(80.0, fmax(-80.0, raw));
    const double e = exp(arg);
    value = 1.0 / (1.0 + e);
    dvalue = (raw < -80.0 || raw > 80.0) ? 0.0 : -16.0 * rc * e / (r * r * ((1.0 + e) * (1.0 + e)));
}

__device__ __forceinline__ double dist2(const double* x, int a, int b) {
    const double dx = x[3 * b] - x[3 * a], dy = x[3 * b + 1] - x[3 * a + 1], dz = x[3 * b + 2] - x[3 * a + 2];
    return dx * dx + dy * dy + dz * dz;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Lattice images of a periodic instance within a cutoff: every translation
// T = n0 a + n1 b + n2 c with |n_k| <= ceil(cutoff / h_k) + 1 (lattice.rs:249-261);
// the callers test the distances themselves.
__device__ __forceinline__ void image_ranges(const CellRec& c, double cut, int r[3]) {
    for (int k = 0; k < 3; ++k) r[k] = (int)ceil(cut / c.h[k]) + 1;
}
__device__ __forceinline__ void image_vector(const CellRec& c, int n0, int n1, int n2, double t[3]) {
    for (int k = 0; k < 3; ++k) t[k] = n0 * c.a[0][k] + n1 * c.a[1][k] + n2 * c.a[2][k];
}
// Overlap screen of the periodic image sums (basis.rs:373-380, applied in
// pbc/bloch.rs:140-146 and pbc/gradient.rs:320): an atom pair whose most diffuse
// primitives overlap below exp(-40) is skipped.
__device__ __forceinline__ bool overlap_screened(const DeviceTables& t, int za, int zb, double r2) {
    const double ea = t.elements[za].min_exponent, eb = t.elements[zb].min_exponent;
    return r2 * ea * eb > 40.0 * (ea + eb);
}

// Deterministic block sum (blockDim.x a multiple of 64, <= 1024); result on every thread.
__device__ double block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += lds[w];
    __syncthreads();
    return s;
}

}  // namespace

// ---------------------------------------------------------------------------
// Coordination numbers for H0 (cutoff cn_cut) and for D3 (25 bohr), plus the
// D3 Gaussian reference weights (coordination.rs:51-94; dispersion.rs:1629-1692).
// D3 reference weights of atom a and their CN derivatives from its D3
// coordination number cd (dispersion.rs:1650-1700), written to gw / dgw.
__device__ void d3_weights(const DeviceTables& t, int z, int a, double cd, double* gw, double* dgw) {
    const ElementType& e = t.elements[z];
    const int nref = e.d3_nref;
    double mx = -INFINITY;
    for (int r = 0; r < nref; ++r) {
        const double d = cd - e.d3_refcn[r];
        mx = fmax(mx, -4.0 * d * d);
    }
    double raw[kMaxRef], norm = 0.0, dnorm = 0.0;
    for (int r = 0; r < nref; ++r) {
        const double d = cd - e.d3_refcn[r];
        raw[r] = exp(-4.0 * d * d - mx);
        norm += raw[r];
        dnorm += 8.0 * (e.d3_refcn[r] - cd) * raw[r];
    }
    const double ni = 1.0 / norm;
    for (int r = 0; r < kMaxRef; ++r) {
        double w = 0.0, dw = 0.0;
        if (r < nref) {
            const double expd = 8.0 * (e.d3_refcn[r] - cd) * raw[r];
            w = raw[r] * ni;
            dw = expd * ni - raw[r] * dnorm * ni * ni;
            if (!isfinite(w)) w = 0.0;
            if (!isfinite(dw)) dw = 0.0;
        }
        gw[a * kMaxRef + r] = w;
        dgw[a * kMaxRef + r] = dw;
    }
}

__global__ void k_coordination(DeviceBatch b, DeviceTables t, double cn_cut, int cn_on, double* cn, double* cn_d3,
                               double* gw, double* dgw, int natoms) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms) return;
    const int inst = b.atom_inst[a];
    // periodic atoms: k_coordination_pbc
    if (periodic_instance(b, inst)) return;
    const int a0 = b.atom0[inst], a1 = b.atom0[inst + 1];
    const double* x = b.atom_x;
    const double rca = t.elements[b.atom_z[a]].covalent_radius;
    double c = 0.0, cd = 0.0;
    const double cc2 = cn_cut * cn_cut, cd2 = 25.0 * 25.0;
    for (int o = a0; o < a1; ++o) {
        if (o == a) continue;
        const double r2 = dist2(x, a, o);
        if (r2 <= kPairEps2) continue;
        const double rc = rca + t.elements[b.atom_z[o]].covalent_radius;
        if (rc <= 1e-12) continue;
        const double r = sqrt(r2);
        double v, dv;
        if (r2 <= cc2 || r2 <= cd2) count_fn(r, rc, v, dv);
        if (r2 <= cc2) c += v;
        if (r2 <= cd2) cd += v;
    }
    cn[a] = cn_on ? c : 0.0;
    cn_d3[a] = cd;
    d3_weights(t, b.atom_z[a], a, cd, gw, dgw);
}

// Lattice-summed coordination numbers of the listed periodic atoms: every
// (atom, image) other than A itself in the origin cell, so a pair with its own
// images counts from both sides (coordination.rs:64-89 over
// pairlist.rs:244-317).  One 64-lane work-group per atom; a work item is one
// (plane n0 of images, atom o) pair, o fastest over the lanes, that walks the
// images of its plane; the lanes' sums are reduced in a fixed order.
__global__ void __launch_bounds__(64) k_coordination_pbc(DeviceBatch b, DeviceTables t, const int* patoms,
                                                         double cn_cut, int cn_on, double* cn, double* cn_d3,
                                                         double* gw, double* dgw) {
    const int a = patoms[blockIdx.x];
    const int inst = b.atom_inst[a];
    const int a0 = b.atom0[inst], nat = b.atom0[inst + 1] - a0;
    const double* x = b.atom_x;
    const double rca = t.elements[b.atom_z[a]].covalent_radius;
    double c = 0.0, cd = 0.0;
    const double cc2 = cn_cut * cn_cut, cd2 = 25.0 * 25.0;
    const CellRec& cr = b.cell[inst];
    int rg[3];
    image_ranges(cr, fmax(cn_cut, 25.0), rg);
    for (int w = threadIdx.x; w < nat * (2 * rg[0] + 1); w += 64) {
        const int n0 = w / nat - rg[0], o = a0 + w % nat;
        const double rc = rca + t.elements[b.atom_z[o]].covalent_radius;
        if (rc <= 1e-12) continue;
        for (int n1 = -rg[1]; n1 <= rg[1]; ++n1)
            for (int n2 = -rg[2]; n2 <= rg[2]; ++n2) {
                if (o == a && n0 == 0 && n1 == 0 && n2 == 0) continue;
                double tv[3];
                image_vector(cr, n0, n1, n2, tv);
                const double dx = x[3 * o] + tv[0] - x[3 * a], dy = x[3 * o + 1] + tv[1] - x[3 * a + 1],
                             dz = x[3 * o + 2] + tv[2] - x[3 * a + 2];
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 <= kPairEps2 || (r2 > cc2 && r2 > cd2)) continue;
                double v, dv;
                count_fn(sqrt(r2), rc, v, dv);
                if (r2 <= cc2) c += v;
                if (r2 <= cd2) cd += v;
            }
    }
    c = wave_sum(c);
    cd = wave_sum(cd);
    if (threadIdx.x != 0) return;
    cn[a] = cn_on ? c : 0.0;
    cn_d3[a] = cd;
    d3_weights(t, b.atom_z[a], a, cd, gw, dgw);
}

// (S/H0 kernel: see the shell-class kernels below)

// Diagonal of the padded rows/columns: S = 1, H0 = F = kPadEnergy.
__global__ void k_pad_diagonal(DeviceBatch b, const int* act) {
    const int inst = act[blockIdx.x];
    const int n = b.nao[inst], ld = b.ld[inst];
    for (int i = n + threadIdx.x; i < ld; i += blockDim.x) {
        const size_t idx = (size_t)i * ld + i;
        ((gptr<double>)b.S[inst])[idx] = 1.0;
        ((gptr<double>)b.H0[inst])[idx] = kPadEnergy;
        ((gptr<double>)b.F[inst])[idx] = kPadEnergy;
    }
}

__global__ void k_mark_failed(const int* act, int n, const int* flag, SccState* st) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n && flag[j] != 0 && st[act[j]].status == 0) st[act[j]].status = 2;
}

// Cb = (L^-1)^T from the batched triangular inverse of the Cholesky factor
// (stored column-major lower in Linv); also checks the factor's pivots.
__global__ void k_make_cb(DeviceBatch b, double* const* Linv, const int* act, double eig_tol, int* fail) {
    const int inst = act[blockIdx.y];
    const int ld = b.ld[inst];
    const int n = b.nao[inst];
    gptr<double> cb = (gptr<double>)b.CB0[inst];
    gptr<const double> li = (gptr<const double>)Linv[blockIdx.y];
    const size_t total = (size_t)ld * ld;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (size_t)gridDim.x * blockDim.x) {
        const int row = (int)(k % ld), col = (int)(k / ld);
        // Cb(row, col) = Linv(col, row) for col >= row
        cb[k] = col >= row ? li[(size_t)row * ld + col] : 0.0;
        if (row == col && row < n) {
            // L_ii = 1/Linv_ii; an overlap eigenvalue <= tol forces a pivot^2 <= tol
            const double lii = 1.0 / li[(size_t)row * ld + row];
            if (!(lii * lii > eig_tol)) fail[inst] = 1;
        }
    }
}

// ---------------------------------------------------------------------------
// Shell potential v = gamma q (+ on-site third order) (coulomb.rs:82-107, :140-193).
// vg gets gamma q only; vt the full potential.
// Instances up to this many shells stage their shells' positions, inverse
// hardnesses and charges in LDS for k_shell_potential (24 KB).
constexpr int kShellLds = 512;

__global__ void k_shell_potential(DeviceBatch b, DeviceTables t, const int* act, const double* q, double* vg,
                                  double* vt) {
    __shared__ double sx[kShellLds], sy[kShellLds], sz[kShellLds], sih[kShellLds], sq[kShellLds];
    __shared__ int sa[kShellLds];
    const int inst = act[blockIdx.y];
    if (inst < 0) return;  // gated off in a captured iteration
    const int s0 = b.sh0[inst], nsh = b.nsh[inst];
    const double* x = b.atom_x;
    const bool pbc = periodic_instance(b, inst);
    const bool staged = !pbc && nsh <= kShellLds;  // uniform over the work-group
    if (staged) {
        for (int j = threadIdx.x; j < nsh; j += blockDim.x) {
            const int sj = s0 + j, aj = b.sh_atom[sj];
            sa[j] = aj;
            sx[j] = x[3 * aj];
            sy[j] = x[3 * aj + 1];
            sz[j] = x[3 * aj + 2];
            sih[j] = 1.0 / t.shells[b.sh_type[sj]].hardness;
            sq[j] = q[sj];
        }
        __syncthreads();
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsh) return;
    const int si = s0 + i;
    const int ai = b.sh_atom[si];
    const double hi = t.shells[b.sh_type[si]].hardness;
    double v = 0.0;
    if (pbc) {
        // the Ewald-summed matrix of the periodic instance (pbc/ewald.rs:253-320)
        const double* g = b.gam + b.cell[inst].gam0 + (size_t)i * nsh;
        for (int j = 0; j < nsh; ++j) v += g[j] * q[s0 + j];
    } else if (staged) {
        // the same operations as the global-memory loop below, operands from LDS
        const double ihi = 1.0 / hi;
        const double xi = x[3 * ai], yi = x[3 * ai + 1], zi = x[3 * ai + 2];
        for (int j = 0; j < nsh; ++j) {
            const double h = 2.0 / (ihi + sih[j]);
            double gam;
            if (ai == sa[j]) {
                gam = h;
            } else {
                const double dx = sx[j] - xi, dy = sy[j] - yi, dz = sz[j] - zi;
                gam = 1.0 / sqrt(dx * dx + dy * dy + dz * dz + 1.0 / (h * h));
            }
            v += gam * sq[j];
        }
    } else {
        for (int j = 0; j < nsh; ++j) {
            const int sj = s0 + j;
            const int aj = b.sh_atom[sj];
            const double hj = t.shells[b.sh_type[sj]].hardness;
            const double h = 2.0 / (1.0 / hi + 1.0 / hj);
            double gam;
            if (ai == aj) gam = h;
            else gam = 1.0 / sqrt(dist2(x, ai, aj) + 1.0 / (h * h));
            v += gam * q[sj];
        }
    }
    vg[si] = v;
    // third order: Gamma_A q_A^2 on every shell of atom A
    const int f0 = b.atom_sh0[ai];
    const int za = b.atom_z[ai];
    const int nsa = t.elements[za].nshell;
    double qa = 0.0;
    for (int k = 0; k < nsa; ++k) qa += q[f0 + k];
    vt[si] = v + qa * qa * t.elements[za].gam3;
}

// F = H0 - 1/2 S (v_mu + v_nu) (electronic.rs:3133-3158); one block per column.
__global__ void k_fock(DeviceBatch b, const int* act, const double* vt) {
    const int inst = act[blockIdx.y];
    if (inst < 0) return;
    const int n = b.nao[inst], ld = b.ld[inst];
    const int col = blockIdx.x;
    if (col >= n) return;
    const int a0 = b.ao0[inst];
    const double vc = vt[b.ao_sh[a0 + col]];
    gptr<const double> S = (gptr<const double>)b.S[inst] + (size_t)col * ld;
    gptr<const double> H = (gptr<const double>)b.H0[inst] + (size_t)col * ld;
    gptr<double> F = (gptr<double>)b.F[inst] + (size_t)col * ld;
    for (int row = threadIdx.x; row < n; row += blockDim.x) {
        const double vr = vt[b.ao_sh[a0 + row]];
        F[row] = H[row] + (-S[row] * 0.5 * (vr + vc));
    }
}

// Lower triangle of A = Cb^T F Cb without forming F.  With F = H0 - (V S + S V)/2
// (electronic.rs:3133-3158, V = diag of the AO potential), S = L L^T and
// Cb = L^-T:
//     A = A0 - (M + M^T) / 2,   A0 = Cb^T H0 Cb,   M = L^-1 V L  (lower triangular),
// so A_ij = A0_ij - M_ij / 2 below the diagonal and A_ii = A0_ii - v_i.  M costs
// n^3 / 3 flops against 4 n^3 for the two GEMMs of the explicit congruence.  One
// 64 x 64 tile of the lower triangle per work-group (256 threads, 4 x 4 outputs
// each); M_ij = sum_{k=j..i} Linv_ik v_k L_kj runs over 16-row chunks in LDS.
__global__ void __launch_bounds__(256) k_reduce_fock(DeviceBatch b, const int* act, const double* vt,
                                                     double* const* Cbp, double* const* Lp, double* const* A0p,
                                                     double* const* Ap) {
    const int ti = blockIdx.x, tj = blockIdx.y;
    if (tj > ti) return;
    const int inst = act[blockIdx.z];
    if (inst < 0) return;
    const int n = b.nao[inst], ld = b.ld[inst];
    const int I0 = ti * 64, J0 = tj * 64;
    if (I0 >= ld) return;
    const int a0 = b.ao0[inst];
    gptr<const double> Cb = (gptr<const double>)Cbp[blockIdx.z];
    gptr<const double> L = (gptr<const double>)Lp[blockIdx.z];
    gptr<const double> A0 = (gptr<const double>)A0p[blockIdx.z];
    gptr<double> A = (gptr<double>)Ap[blockIdx.z];
    __shared__ double Bi[16][64], Bj[16][64];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[p][c] = 0.0;
    const int kend = min(I0 + 64, ld);
    for (int k0 = J0; k0 < kend; k0 += 16) {
        for (int q = threadIdx.x; q < 16 * 64; q += 256) {
            const int kk = q & 15, ii = q >> 4;
            const int k = k0 + kk, i = I0 + ii, j = J0 + ii;
            const double vk = k < n ? vt[b.ao_sh[a0 + k]] : 0.0;
            Bi[kk][ii] = (k < ld && i < ld && k <= i) ? Cb[k + (size_t)i * ld] * vk : 0.0;
            Bj[kk][ii] = (k < ld && j < ld && k >= j) ? L[k + (size_t)j * ld] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            double ai[4], bj[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) ai[p] = Bi[kk][ty * 4 + p];
#pragma unroll
            for (int c = 0; c < 4; ++c) bj[c] = Bj[kk][tx * 4 + c];
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[p][c] += ai[p] * bj[c];
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int i = I0 + ty * 4 + p, j = J0 + tx * 4 + c;
            if (i < ld && j < ld && i >= j) {
                const size_t o = i + (size_t)j * ld;
                if (i == j) A[o] = A0[o] - (i < n ? vt[b.ao_sh[a0 + i]] : 0.0);
                else A[o] = A0[o] - 0.5 * acc[p][c];
            }
        }
}

// F += b (S - 1/2 S P S) for instances on a level-shift rung (electronic.rs:2120-2133).
__global__ void k_level_shift(DeviceBatch b, const int* act, const double* shift, double* const* SPS) {
    const int inst = act[blockIdx.y];
    const int n = b.nao[inst], ld = b.ld[inst];
    const int col = blockIdx.x;
    if (col >= n) return;
    const double ls = shift[blockIdx.y];
    gptr<const double> S = (gptr<const double>)b.S[inst] + (size_t)col * ld;
    gptr<const double> M = (gptr<const double>)SPS[blockIdx.y] + (size_t)col * ld;
    gptr<double> F = (gptr<double>)b.F[inst] + (size_t)col * ld;
    for (int row = threadIdx.x; row < n; row += blockDim.x) F[row] += ls * (S[row] - 0.5 * M[row]);
}

// ---------------------------------------------------------------------------
// Fermi occupations by bisection to the f64 fixed point (electronic.rs:3853-3921).
// eig: eigenvalues of the active list (stride ld), copied into the per-instance
// orbital-energy vector.  Padded orbitals get zero occupation.
__global__ void k_fermi(DeviceBatch b, const int* act, const double* eig, int eig_stride, double* eps_all,
                        double* occ_all, SccState* st) {
    __shared__ double lds[16];
    const int inst = act[blockIdx.x];
    const int n = b.nao[inst], ld = b.ld[inst];
    const double* e = eig + (size_t)blockIdx.x * eig_stride;
    double* eps = eps_all + b.vec0[inst];
    double* occ = occ_all + b.vec0[inst];
    SccState& s = st[inst];
    const double kt = s.kt, nelec = s.nelec;
    for (int i = threadIdx.x; i < ld; i += blockDim.x) eps[i] = e[i];
    double mn = INFINITY, mx = -INFINITY;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        mn = fmin(mn, e[i]);
        mx = fmax(mx, e[i]);
    }
    // block min / max
    for (int off = 32; off > 0; off >>= 1) {
        mn = fmin(mn, __shfl_down(mn, off, 64));
        mx = fmax(mx, __shfl_down(mx, off, 64));
    }
    __shared__ double smn[16], smx[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) { smn[wid] = mn; smx[wid] = mx; }
    __syncthreads();
    mn = smn[0];
    mx = smx[0];
    for (int w = 1; w < nw; ++w) { mn = fmin(mn, smn[w]); mx = fmax(mx, smx[w]); }
    __syncthreads();
    double mu = 0.0, ts = 0.0;
    int kneed = 0;
    if (kt <= 0.0) {
        // aufbau (electronic.rs:3829-3851)
        if (threadIdx.x == 0) {
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
            mu = n ? e[homo] : 0.0;
            kneed = homo + 1;
        }
    } else if (nelec <= 0.0 || nelec >= 2.0 * n) {
        const double f = nelec <= 0.0 ? 0.0 : 2.0;
        for (int i = threadIdx.x; i < ld; i += blockDim.x) occ[i] = i < n ? f : 0.0;
        mu = nelec <= 0.0 ? e[0] : e[n - 1];
        kneed = nelec <= 0.0 ? 0 : n;
    } else {
        double lo = mn - 100.0 * kt - 10.0, hi = mx + 100.0 * kt + 10.0;
        for (int it = 0; it < 200; ++it) {
            const double m = 0.5 * (lo + hi);
            double part = 0.0;
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                const double xx = fmin(80.0, fmax(-80.0, (e[i] - m) / kt));
                part += 2.0 / (1.0 + exp(xx));
            }
            const double sum = block_sum(part, lds);
            double nlo = lo, nhi = hi;
            if (sum < nelec) nlo = m;
            else nhi = m;
            if (nlo == lo && nhi == hi) break;
            lo = nlo;
            hi = nhi;
        }
        mu = 0.5 * (lo + hi);
        double tpart = 0.0;
        int kpart = 0;
        for (int i = threadIdx.x; i < ld; i += blockDim.x) {
            double f = 0.0;
            if (i < n) {
                const double xx = fmin(80.0, fmax(-80.0, (e[i] - mu) / kt));
                f = 2.0 / (1.0 + exp(xx));
                const double h = fmin(1.0 - 1e-16, fmax(1e-16, 0.5 * f));
                tpart += 2.0 * kt * (h * log(h) + (1.0 - h) * log(1.0 - h));
                if (f >= 1e-28) kpart = max(kpart, i + 1);
            }
            occ[i] = f;
        }
        ts = block_sum(tpart, lds);
        // block max of kpart
        int km = kpart;
        for (int off = 32; off > 0; off >>= 1) km = max(km, __shfl_down(km, off, 64));
        __shared__ int skm[16];
        if (lane == 0) skm[wid] = km;
        __syncthreads();
        km = skm[0];
        for (int w = 1; w < nw; ++w) km = max(km, skm[w]);
        kneed = km;
    }
    if (threadIdx.x == 0) {
        s.mu = mu;
        s.ts = ts;
        s.kneed = kneed;
    }
}

// Mixed precision operands of M = C_k diag(w) C_k^T: C32 = fp32(C_k) and
// A32 = fp32(C_k diag(w)), w = f (density) or eps f (energy-weighted density);
// column col of each at col * ld (rows n..ld-1 zero).
__global__ void k_operands_fp32(DeviceBatch b, const int* act, double* const* Cp, double* const* A32p,
                                double* const* C32p, const double* eps_all, const double* occ_all, int k,
                                int weighted) {
    const int inst = act[blockIdx.y];
    const int n = b.nao[inst], ld = b.ld[inst];
    const int col = blockIdx.x;
    if (col >= k) return;
    const double f = occ_all[b.vec0[inst] + col];
    const double w = weighted ? eps_all[b.vec0[inst] + col] * f : f;
    gptr<const double> c = (gptr<const double>)Cp[blockIdx.y] + (size_t)col * ld;
    gptr<float> a32 = (gptr<float>)A32p[blockIdx.y] + (size_t)col * ld;
    gptr<float> c32 = (gptr<float>)C32p[blockIdx.y] + (size_t)col * ld;
    for (int row = threadIdx.x; row < ld; row += blockDim.x) {
        const double v = row < n ? c[row] : 0.0;
        c32[row] = (float)v;
        a32[row] = (float)(v * w);
    }
}

// M = (X + X^T) / 2 in FP64 from an FP32 product X (n x n, leading dimension ld).
__global__ void k_symmetrize_fp32(DeviceBatch b, const int* act, double* const* Xp, double* const* Mp) {
    __shared__ float t1[32][33], t2[32][33];
    const int inst = act[blockIdx.z];
    const int n = b.nao[inst], ld = b.ld[inst];
    const int bi = blockIdx.x, bj = blockIdx.y;
    if (bj > bi) return;
    if (bi * 32 >= n) return;
    gptr<const float> X = (gptr<const float>)Xp[blockIdx.z];
    gptr<double> M = (gptr<double>)Mp[blockIdx.z];
    const int tx = threadIdx.x & 31, ty0 = threadIdx.x >> 5;
    for (int ty = ty0; ty < 32; ty += 8) {
        const int r1 = bi * 32 + tx, c1 = bj * 32 + ty;
        const int r2 = bj * 32 + tx, c2 = bi * 32 + ty;
        t1[ty][tx] = (r1 < n && c1 < n) ? X[(size_t)c1 * ld + r1] : 0.0f;
        t2[ty][tx] = (r2 < n && c2 < n) ? X[(size_t)c2 * ld + r2] : 0.0f;
    }
    __syncthreads();
    for (int ty = ty0; ty < 32; ty += 8) {
        const int r1 = bi * 32 + tx, c1 = bj * 32 + ty;
        const int r2 = bj * 32 + tx, c2 = bi * 32 + ty;
        if (r1 < n && c1 < n) M[(size_t)c1 * ld + r1] = 0.5 * ((double)t1[ty][tx] + (double)t2[tx][ty]);
        if (bi != bj && r2 < n && c2 < n) M[(size_t)c2 * ld + r2] = 0.5 * ((double)t2[ty][tx] + (double)t1[tx][ty]);
    }
}

// Large matrices: the per-64-column-block row sums of k_density added up in
// block order (eight interleaved partial sums, fixed combination) into block 0,
// one thread per AO, so that k_qout (blocks = 1) does not walk every block.
__global__ void k_collapse_rowsums(DeviceBatch b, const int* act, double* rowS, double* rowH, int stride) {
    const int inst = act[blockIdx.y];
    if (inst < 0) return;
    const int a0 = b.ao0[inst], n = b.nao[inst];
    const int mu = blockIdx.x * blockDim.x + threadIdx.x;
    if (mu >= n) return;
    const int nblk = (n + 63) / 64;
    for (double* part : {rowS, rowH}) {
        double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        int blk = 0;
        for (; blk + 8 <= nblk; blk += 8)
            for (int u = 0; u < 8; ++u) acc[u] += part[(size_t)(blk + u) * stride + a0 + mu];
        for (int u = 0; blk < nblk; ++blk, ++u) acc[u] += part[(size_t)blk * stride + a0 + mu];
        part[a0 + mu] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    }
}

// Output shell charges q = n0 - sum_{mu in shell} (S P)_mu mu and the band energy
// (electronic.rs:3160-3188), from the per-64-column-block partial sums
// rowS[blk * stride + ao0 + mu] (summed in block order).
__global__ void k_qout(DeviceBatch b, DeviceTables t, const int* act, const double* rowS, const double* rowH,
                       int stride, int blocks, double* qout, SccState* st) {
    __shared__ double lds[16];
    const int inst = act[blockIdx.x];
    if (inst < 0) return;
    const int s0 = b.sh0[inst], nsh = b.nsh[inst];
    const int a0 = b.ao0[inst], n = b.nao[inst];
    const int nblk = blocks > 0 ? blocks : (n + 63) / 64;
    auto total = [&](const double* part, int mu) {
        double v = 0.0;
        for (int blk = 0; blk < nblk; ++blk) v += part[(size_t)blk * stride + a0 + mu];
        return v;
    };
    for (int i = threadIdx.x; i < nsh; i += blockDim.x) {
        const int si = s0 + i;
        const ShellType& T = t.shells[b.sh_type[si]];
        double pop = 0.0;
        for (int k = 0; k < T.nao; ++k) pop += total(rowS, b.sh_ao[si] + k);
        qout[si] = T.ref_occ - pop;
    }
    double e = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) e += total(rowH, i);
    e = block_sum(e, lds);
    if (threadIdx.x == 0) st[inst].e_el = e;
}

// ---------------------------------------------------------------------------
// Convergence test and charge mixing (electronic.rs:1843-1947, 3203-3311).
struct LadderRung {
    int broyden;
    double mixing, level_shift, temperature;
    int max_iter;
};

// Gauss-Jordan with partial pivoting (electronic.rs:3767-3806) by the whole
// work-group: the pivot search is a block arg-max (first index among equal
// magnitudes, as in the serial loop), each thread owns whole rows in the
// elimination, and every element sees the same operations in the same order as
// in the serial algorithm.  a (m x m, row stride lda) and bvec may be in global
// memory.  Returns the same value in every thread.
__device__ bool solve_gj_block(double* a, double* bvec, int m, int lda, double* rv, int* ri) {
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
    for (int col = 0; col < m; ++col) {
        double best = -1.0;
        int bi = m;
        for (int r = col + tid; r < m; r += nt) {
            const double v = fabs(a[r * lda + col]);
            if (v > best) {
                best = v;
                bi = r;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_down(best, off, 64);
            const int oi = __shfl_down(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) {
                best = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            rv[wid] = best;
            ri[wid] = bi;
        }
        __syncthreads();
        best = rv[0];
        bi = ri[0];
        for (int w = 1; w < nw; ++w)
            if (rv[w] > best || (rv[w] == best && ri[w] < bi)) {
                best = rv[w];
                bi = ri[w];
            }
        __syncthreads();
        if (best < 1e-18) return false;
        const int piv = bi;
        if (piv != col) {
            for (int j = tid; j < m; j += nt) {
                const double t = a[piv * lda + j];
                a[piv * lda + j] = a[col * lda + j];
                a[col * lda + j] = t;
            }
            if (tid == 0) {
                const double t = bvec[piv];
                bvec[piv] = bvec[col];
                bvec[col] = t;
            }
            __syncthreads();
        }
        const double dg = a[col * lda + col];
        __syncthreads();
        for (int j = col + tid; j < m; j += nt) a[col * lda + j] /= dg;
        if (tid == 0) bvec[col] /= dg;
        __syncthreads();
        for (int r = tid; r < m; r += nt) {
            if (r == col) continue;
            const double f = a[r * lda + col];
            if (f == 0.0) continue;
            for (int j = col; j < m; ++j) a[r * lda + j] -= f * a[col * lda + j];
            bvec[r] -= f * bvec[col];
        }
        __syncthreads();
    }
    return true;
}

__global__ void k_scc_update(DeviceBatch b, DeviceTables t, const int* act, SccParams prm, const double* qout,
                             const double* vg_out, double* qin, const double* q0, SccState* st, double* br_df,
                             double* br_u, double* br_omega, double* br_qlast, double* br_dqlast,
                             const int64_t* br_off, double* beta_scratch, int beta_ld) {
    __shared__ double lds[16];
    __shared__ double sh_beta[32 * 33];
    __shared__ double sh_c[32];
    __shared__ int sh_ri[16];
    const int inst = act[blockIdx.x];
    if (inst < 0) return;
    SccState& s = st[inst];
    if (s.status != 0) return;
    // Every thread reads the state fields it needs here, before the first
    // barrier; thread 0 writes the state only after that barrier.  (A field read
    // by the other wavefronts after thread 0 has written it would split the
    // work-group between the branches below.)
    const double s_e_el = s.e_el, s_ts = s.ts, s_e_last = s.e_last;
    const int s_have_energy = s.have_energy, s_rung = s.rung, s_iter = s.iter, s_broyden_iter = s.broyden_iter;
    const int s0 = b.sh0[inst], nsh = b.nsh[inst];
    const int atom0 = b.atom0[inst], nat = b.atom0[inst + 1] - atom0;
    // E2 = 1/2 q.gamma q, E3 = sum_A Gamma_A q_A^3 / 3
    double e2 = 0.0, ss = 0.0;
    for (int i = threadIdx.x; i < nsh; i += blockDim.x) {
        e2 += qout[s0 + i] * vg_out[s0 + i];
        const double d = qout[s0 + i] - qin[s0 + i];
        ss += d * d;
    }
    double e3 = 0.0;
    for (int a = threadIdx.x; a < nat; a += blockDim.x) {
        const int ga = atom0 + a;
        const int f0 = b.atom_sh0[ga];
        const int za = b.atom_z[ga];
        double qa = 0.0;
        for (int k = 0; k < t.elements[za].nshell; ++k) qa += qout[f0 + k];
        e3 += qa * qa * qa * t.elements[za].gam3 / 3.0;
    }
    e2 = 0.5 * block_sum(e2, lds);
    e3 = block_sum(e3, lds);
    ss = block_sum(ss, lds);
    const double rms = nsh ? sqrt(ss / nsh) : 0.0;
    const double energy = s_e_el + e2 + e3 + s_ts;
    const double de = s_have_energy ? fabs(energy - s_e_last) : INFINITY;
    __syncthreads();
    if (threadIdx.x == 0) {
        s.e2 = e2;
        s.e3 = e3;
        s.rms = rms;
        s.energy = energy;
        s.iter += 1;
        s.total_iter += 1;
    }
    if (de < prm.energy_tolerance && rms < prm.charge_tolerance) {
        if (threadIdx.x == 0) s.status = 1;
        return;
    }
    // ---- next input charges
    // a periodic instance runs the caller's scheme once, without level shift
    // and without the fallback ladder (pbc/scf.rs:358-529)
    const bool pbc = periodic_instance(b, inst);
    const LadderRung rung = {prm.rung_broyden[s_rung], prm.rung_mixing[s_rung],
                             pbc ? 0.0 : prm.rung_level_shift[s_rung], prm.rung_temperature[s_rung],
                             prm.rung_max_iter[s_rung]};
    const double mix = rung.mixing;
    double* qi = qin + s0;
    const double* qo = qout + s0;
    bool use_damped = true;
    if (rung.broyden) {
        // BroydenMixer::next; state lives in global memory
        const int mem = prm.broyden_memory;
        const int64_t off = br_off[inst];
        double* df = br_df + off;
        double* u = br_u + off;
        double* omega = br_omega + (size_t)inst * prm.omega_stride;
        double* qlast = br_qlast + s0;
        double* dqlast = br_dqlast + s0;
        const int iter = s_broyden_iter + 1;
        if (iter == 1) {
            for (int i = threadIdx.x; i < nsh; i += blockDim.x) {
                const double dq = qo[i] - qi[i];
                dqlast[i] = dq;
                qlast[i] = qi[i];
            }
            // the damped first step is the candidate; trusted-check below
            use_damped = true;
        } else {
            const int itn = iter - 1;
            const int it1 = (itn - 1) % mem;
            double dqn = 0.0, dfn = 0.0;
            for (int i = threadIdx.x; i < nsh; i += blockDim.x) {
                const double dq = qo[i] - qi[i];
                const double dfc = dq - dqlast[i];
                dqn += dq * dq;
                dfn += dfc * dfc;
            }
            dqn = sqrt(block_sum(dqn, lds));
            dfn = block_sum(dfn, lds);
            const double wfac = 0.01, maxw = 100000.0, minw = 1.0, omega0 = 0.01;
            double om = dqn > wfac / maxw ? wfac / dqn : maxw;
            if (om < minw) om = minw;
            const double inv = 1.0 / fmax(sqrt(dfn), 2.220446049250313e-16);
            for (int i = threadIdx.x; i < nsh; i += blockDim.x) {
                const double dq = qo[i] - qi[i];
                const double dfi = (dq - dqlast[i]) * inv;
                df[(size_t)it1 * nsh + i] = dfi;
                u[(size_t)it1 * nsh + i] = mix * dfi + inv * (qi[i] - qlast[i]);
            }
            if (threadIdx.x == 0) omega[it1] = om;
            __syncthreads();
            const int start = itn > mem ? itn - mem + 1 : 1;
            const int m = itn - start + 1;
            // m <= 32 lives in LDS; longer histories use a caller-provided scratch
            // of (beta_ld + 1) * beta_ld doubles per active instance
            double* scratch = beta_scratch + (size_t)blockIdx.x * (beta_ld + 1) * beta_ld;
            double* beta = m <= kBroydenLdsHistory ? sh_beta : scratch;
            const int bl = m <= kBroydenLdsHistory ? 33 : beta_ld;
            double* cvec = m <= kBroydenLdsHistory ? sh_c : scratch + (size_t)beta_ld * beta_ld;
            // beta and c by dot products; one (r, c) pair per thread-iteration
            for (int pr = threadIdx.x; pr < m * (m + 1); pr += blockDim.x) {
                const int r = pr / (m + 1), c = pr % (m + 1);
                const int ir = (start + r - 1) % mem;
                double dsum = 0.0;
                if (c < m) {
                    const int ic = (start + c - 1) % mem;
                    for (int i = 0; i < nsh; ++i) dsum += df[(size_t)ir * nsh + i] * df[(size_t)ic * nsh + i];
                    double v = omega[ir] * omega[ic] * dsum;
                    if (r == c) v += omega0 * omega0;
                    beta[r * bl + c] = v;
                } else {
                    for (int i = 0; i < nsh; ++i) dsum += df[(size_t)ir * nsh + i] * (qo[i] - qi[i]);
                    cvec[r] = omega[ir] * dsum;
                }
            }
            __syncthreads();
            const bool solved = solve_gj_block(beta, cvec, m, bl, lds, sh_ri);
            if (solved) {
                // candidate = q + damp dq - sum_r omega_r c_r u_r ; then qlast/dqlast update
                int bad = 0;
                for (int i = threadIdx.x; i < nsh; i += blockDim.x) {
                    const double dq = qo[i] - qi[i];
                    double v = qi[i] + mix * dq;
                    for (int r = 0; r < m; ++r) {
                        const int ir = (start + r - 1) % mem;
                        v -= omega[ir] * cvec[r] * u[(size_t)ir * nsh + i];
                    }
                    if (!isfinite(v) || fabs(v) >= 10.0) bad = 1;
                    dqlast[i] = dq;     // BroydenMixer::next updates dqlast/qlast on success
                    const double qold = qi[i];
                    qlast[i] = qold;
                    // write candidate into qin directly; revert below if untrusted
                    qi[i] = v;
                }
                const double nb = block_sum((double)bad, lds);
                if (nb == 0.0) {
                    use_damped = false;
                } else {
                    // revert: qin = qlast (the pre-step charges), then take the damped step
                    for (int i = threadIdx.x; i < nsh; i += blockDim.x) qi[i] = qlast[i];
                    __syncthreads();
                    use_damped = true;
                }
            } else {
                use_damped = true;  // singular: qlast/dqlast unchanged (reference returns early)
            }
        }
        if (threadIdx.x == 0) s.broyden_iter = iter;
    }
    __syncthreads();
    if (use_damped) {
        for (int i = threadIdx.x; i < nsh; i += blockDim.x) qi[i] = qi[i] + mix * (qo[i] - qi[i]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s.have_prev_p = rung.level_shift != 0.0 ? 1 : 0;
        s.e_last = energy;
        s.have_energy = 1;
    }
    // rung exhausted -> next rung or failure (electronic.rs:1413-1446, 1950-1955)
    if (s_iter + 1 >= rung.max_iter) {
        __syncthreads();
        if (s_rung >= 3 || pbc) {
            if (threadIdx.x == 0) s.status = 2;
            return;
        }
        const int nr = s_rung + 1;
        for (int i = threadIdx.x; i < nsh; i += blockDim.x) qi[i] = q0[s0 + i];
        __syncthreads();
        if (threadIdx.x == 0) {
            s.rung = nr;
            s.iter = 0;
            s.have_energy = 0;
            s.have_prev_p = 0;
            s.broyden_iter = 0;
            s.mixing = prm.rung_mixing[nr];
            s.level_shift = prm.rung_level_shift[nr];
            s.temperature = prm.rung_temperature[nr];
            s.kt = fmax(prm.rung_temperature[nr], 0.0) * kBoltzmannHartree;
        }
    }
}

// ---------------------------------------------------------------------------
// Geometry-only terms of molecular atoms.  One thread per atom A, looping over
// the instance's other atoms; every pair is evaluated from both sides so the
// per-atom sums need no atomics and are deterministic.
//   repulsion (repulsion.rs:39-59, :139-159), 25 bohr
//   D3(BJ) two-body energy, gradient and dE/dCN (dispersion.rs:1232-1300)
__global__ void k_geometry_terms(DeviceBatch b, DeviceTables t, const double* gw, const double* cn_d3, int disp_on,
                                 double* e_rep_atom, double* e_disp_atom, double* g_rep, double* g_disp,
                                 double* dedcn_disp, int natoms) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms) return;
    const int inst = b.atom_inst[a];
    // periodic atoms: k_geometry_terms_pbc
    if (periodic_instance(b, inst)) return;
    const int a0 = b.atom0[inst], a1 = b.atom0[inst + 1];
    const double* x = b.atom_x;
    const int za = b.atom_z[a];
    const ElementType& EA = t.elements[za];
    const Globals& g = t.glob;
    double er = 0.0, ed_sum = 0.0, gr[3] = {0, 0, 0}, gd[3] = {0, 0, 0}, dcn = 0.0;
    const double* gwa = gw + (size_t)a * kMaxRef;
    const double* dga = gw + (size_t)natoms * kMaxRef + (size_t)a * kMaxRef;  // dgw follows gw
    for (int o = a0; o < a1; ++o) {
        if (o == a) continue;
        const double r2 = dist2(x, a, o);
        if (r2 <= kPairEps2) continue;
        const double r = sqrt(r2);
        const int zb = b.atom_z[o];
        const ElementType& EB = t.elements[zb];
        const double dr[3] = {x[3 * o] - x[3 * a], x[3 * o + 1] - x[3 * a + 1], x[3 * o + 2] - x[3 * a + 2]};
        if (r2 <= 25.0 * 25.0) {
            const double alpha = sqrt(EA.repa * EB.repa);
            const double zeff = EA.repb * EB.repb;
            const double k = (za <= 2 && zb <= 2) ? g.kexplight : g.kexp;
            const double rk = pow(r, k);
            const double e = zeff * exp(-alpha * rk) / r;
            const double pref = (alpha * rk * k + 1.0) * e / (r * r);
            if (o > a) er += e;
            for (int d = 0; d < 3; ++d) gr[d] += dr[d] * pref;
        }
        if (disp_on && r2 <= 50.0 * 50.0 && r2 > 1e-14) {
            const double r4r2 = 3.0 * EA.d3_r4r2 * EB.d3_r4r2;
            const double r0 = g.a1 * sqrt(r4r2) + g.a2;
            const double r0_2 = r0 * r0, r0_6 = r0_2 * r0_2 * r0_2, r0_8 = r0_6 * r0_2;
            const double r4 = r2 * r2;
            const double t6 = 1.0 / (r4 * r2 + r0_6), t8 = 1.0 / (r4 * r4 + r0_8);
            const double d6 = -6.0 * r4 * t6 * t6, d8 = -8.0 * r4 * r2 * t8 * t8;
            const double ed = g.s6 * t6 + g.s8 * r4r2 * t8;
            const double gdv = g.s6 * d6 + g.s8 * r4r2 * d8;
            // C6_AB and dC6_AB/dCN_A from the reference weights (dispersion.rs:1753-1787)
            const double* gwb = gw + (size_t)o * kMaxRef;
            int hi, lo, ia_is_hi;
            if (za > zb) { hi = za; lo = zb; ia_is_hi = 1; }
            else { hi = zb; lo = za; ia_is_hi = 0; }
            const size_t pair = (size_t)lo + (size_t)hi * (hi - 1) / 2 - 1;
            const double* tab = t.d3_c6 + pair * kMaxRef * kMaxRef;
            double c6 = 0.0, dc6 = 0.0;
            for (int ra = 0; ra < EA.d3_nref; ++ra)
                for (int rb = 0; rb < EB.d3_nref; ++rb) {
                    // table index [ref of lower Z][ref of higher Z]; for equal Z
                    // the A reference is the slow index (dispersion.rs:3166-3172)
                    double ref;
                    if (za == zb) ref = tab[ra * kMaxRef + rb];
                    else ref = ia_is_hi ? tab[rb * kMaxRef + ra] : tab[ra * kMaxRef + rb];
                    c6 += gwa[ra] * gwb[rb] * ref;
                    dc6 += dga[ra] * gwb[rb] * ref;
                }
            if (o > a) ed_sum -= c6 * ed;
            for (int d = 0; d < 3; ++d) gd[d] += (-dr[d]) * (-c6 * gdv);
            dcn -= dc6 * ed;
        }
    }
    e_rep_atom[a] = er;
    e_disp_atom[a] = ed_sum;
    for (int d = 0; d < 3; ++d) {
        g_rep[3 * a + d] = gr[d];
        g_disp[3 * a + d] = gd[d];
    }
    dedcn_disp[a] = dcn;
}

// Repulsion and D3 two-body terms of the listed periodic atoms as image sums
// (repulsion.rs:39-59, dispersion.rs:1258-1295 over pairlist.rs:244-317).
// Every (atom, image) pair is seen from both atoms, so each adds half its
// energy; a pair of an atom with its own image moves nothing and adds no force,
// but its D3 term enters dE/dCN.  One 64-lane work-group per atom; a work item
// is one (plane n0 of images, atom o) pair, o fastest over the lanes, with the
// pair's C6 and damping radii computed once for its plane; the lanes' sums are
// reduced in a fixed order.
__global__ void __launch_bounds__(64) k_geometry_terms_pbc(DeviceBatch b, DeviceTables t, const int* patoms,
                                                           const double* gw, int disp_on, double* e_rep_atom,
                                                           double* e_disp_atom, double* g_rep, double* g_disp,
                                                           double* dedcn_disp, int natoms) {
    const int a = patoms[blockIdx.x];
    const int inst = b.atom_inst[a];
    const int a0 = b.atom0[inst], nat = b.atom0[inst + 1] - a0;
    const double* x = b.atom_x;
    const int za = b.atom_z[a];
    const ElementType& EA = t.elements[za];
    const Globals& g = t.glob;
    double er = 0.0, ed_sum = 0.0, gr[3] = {0, 0, 0}, gd[3] = {0, 0, 0}, dcn = 0.0;
    const double* gwa = gw + (size_t)a * kMaxRef;
    const double* dga = gw + (size_t)natoms * kMaxRef + (size_t)a * kMaxRef;  // dgw follows gw
    const CellRec& cr = b.cell[inst];
    int rg[3];
    image_ranges(cr, disp_on ? 50.0 : 25.0, rg);
    for (int w = threadIdx.x; w < nat * (2 * rg[0] + 1); w += 64) {
        const int n0 = w / nat - rg[0], o = a0 + w % nat;
        const int zb = b.atom_z[o];
        const ElementType& EB = t.elements[zb];
        double c6 = 0.0, dc6 = 0.0;
        if (disp_on) {
            const double* gwb = gw + (size_t)o * kMaxRef;
            int hi, lo, ia_is_hi;
            if (za > zb) { hi = za; lo = zb; ia_is_hi = 1; }
            else { hi = zb; lo = za; ia_is_hi = 0; }
            const size_t pair = (size_t)lo + (size_t)hi * (hi - 1) / 2 - 1;
            const double* tab = t.d3_c6 + pair * kMaxRef * kMaxRef;
            for (int ra = 0; ra < EA.d3_nref; ++ra)
                for (int rb = 0; rb < EB.d3_nref; ++rb) {
                    double ref;
                    if (za == zb) ref = tab[ra * kMaxRef + rb];
                    else ref = ia_is_hi ? tab[rb * kMaxRef + ra] : tab[ra * kMaxRef + rb];
                    c6 += gwa[ra] * gwb[rb] * ref;
                    dc6 += dga[ra] * gwb[rb] * ref;
                }
        }
        const double alpha = sqrt(EA.repa * EB.repa);
        const double zeff = EA.repb * EB.repb;
        const double kx = (za <= 2 && zb <= 2) ? g.kexplight : g.kexp;
        const double r4r2 = 3.0 * EA.d3_r4r2 * EB.d3_r4r2;
        const double r0 = g.a1 * sqrt(r4r2) + g.a2;
        const double r0_2 = r0 * r0, r0_6 = r0_2 * r0_2 * r0_2, r0_8 = r0_6 * r0_2;
        for (int n1 = -rg[1]; n1 <= rg[1]; ++n1)
            for (int n2 = -rg[2]; n2 <= rg[2]; ++n2) {
                if (o == a && n0 == 0 && n1 == 0 && n2 == 0) continue;
                double tv[3];
                image_vector(cr, n0, n1, n2, tv);
                const double dr[3] = {x[3 * o] + tv[0] - x[3 * a], x[3 * o + 1] + tv[1] - x[3 * a + 1],
                                      x[3 * o + 2] + tv[2] - x[3 * a + 2]};
                const double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                if (r2 <= kPairEps2) continue;
                if (r2 <= 25.0 * 25.0) {
                    const double r = sqrt(r2);
                    const double rk = pow(r, kx);
                    const double e = zeff * exp(-alpha * rk) / r;
                    const double pref = (alpha * rk * kx + 1.0) * e / (r * r);
                    er += 0.5 * e;
                    if (o != a)
                        for (int d = 0; d < 3; ++d) gr[d] += dr[d] * pref;
                }
                if (disp_on && r2 <= 50.0 * 50.0 && r2 > 1e-14) {
                    const double r4 = r2 * r2;
                    const double t6 = 1.0 / (r4 * r2 + r0_6), t8 = 1.0 / (r4 * r4 + r0_8);
                    const double d6 = -6.0 * r4 * t6 * t6, d8 = -8.0 * r4 * r2 * t8 * t8;
                    const double ed = g.s6 * t6 + g.s8 * r4r2 * t8;
                    const double gdv = g.s6 * d6 + g.s8 * r4r2 * d8;
                    ed_sum -= 0.5 * c6 * ed;
                    if (o != a)
                        for (int d = 0; d < 3; ++d) gd[d] += (-dr[d]) * (-c6 * gdv);
                    dcn -= dc6 * ed;
                }
            }
    }
    er = wave_sum(er);
    ed_sum = wave_sum(ed_sum);
    dcn = wave_sum(dcn);
    for (int d = 0; d < 3; ++d) {
        gr[d] = wave_sum(gr[d]);
        gd[d] = wave_sum(gd[d]);
    }
    if (threadIdx.x != 0) return;
    e_rep_atom[a] = er;
    e_disp_atom[a] = ed_sum;
    for (int d = 0; d < 3; ++d) {
        g_rep[3 * a + d] = gr[d];
        g_disp[3 * a + d] = gd[d];
    }
    dedcn_disp[a] = dcn;
}

// CN chain rule of a coordination number with cutoff cut:
// g_A += sum_B (dE/dCN_A + dE/dCN_B) dCN/dr (R_A - R_B)/r (gradient.rs:475-489).
__global__ void k_cn_chain(DeviceBatch b, DeviceTables t, const double* dedcn, double cut, double reps, double* g,
                           int natoms) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms) return;
    const int inst = b.atom_inst[a];
    const int a0 = b.atom0[inst], a1 = b.atom0[inst + 1];
    const double* x = b.atom_x;
    const double rca = t.elements[b.atom_z[a]].covalent_radius;
    const double c2 = cut * cut;
    double acc[3] = {0, 0, 0};
    // periodic atoms: k_cn_chain_pbc
    if (periodic_instance(b, inst)) return;
    for (int o = a0; o < a1; ++o) {
        if (o == a) continue;
        const double r2 = dist2(x, a, o);
        if (r2 <= kPairEps2 || r2 > c2) continue;
        const double rc = rca + t.elements[b.atom_z[o]].covalent_radius;
        if (rc <= 1e-12) continue;
        const double r = sqrt(r2);
        if (r <= reps) continue;
        double v, dv;
        count_fn(r, rc, v, dv);
        const double pref = (dedcn[a] + dedcn[o]) * dv / r;
        for (int d = 0; d < 3; ++d) acc[d] += (x[3 * a + d] - x[3 * o + d]) * pref;
    }
    for (int d = 0; d < 3; ++d) g[3 * a + d] += acc[d];
}

// CN chain rule of the listed periodic atoms: every image of every other atom;
// an atom's own images are skipped (gradient.rs:477-479 over the periodic pair
// list).  One 64-lane work-group per atom; a work item is one (plane n0 of
// images, atom o) pair, o fastest over the lanes; fixed reduction order.
__global__ void __launch_bounds__(64) k_cn_chain_pbc(DeviceBatch b, DeviceTables t, const int* patoms,
                                                     const double* dedcn, double cut, double reps, double* g) {
    const int a = patoms[blockIdx.x];
    const int inst = b.atom_inst[a];
    const int a0 = b.atom0[inst], nat = b.atom0[inst + 1] - a0;
    const double* x = b.atom_x;
    const double rca = t.elements[b.atom_z[a]].covalent_radius;
    const double c2 = cut * cut;
    double acc[3] = {0, 0, 0};
    const CellRec& cr = b.cell[inst];
    int rg[3];
    image_ranges(cr, cut, rg);
    for (int w = threadIdx.x; w < nat * (2 * rg[0] + 1); w += 64) {
        const int n0 = w / nat - rg[0], o = a0 + w % nat;
        if (o == a) continue;
        const double rc = rca + t.elements[b.atom_z[o]].covalent_radius;
        if (rc <= 1e-12) continue;
        const double dedcn_ao = dedcn[a] + dedcn[o];
        for (int n1 = -rg[1]; n1 <= rg[1]; ++n1)
            for (int n2 = -rg[2]; n2 <= rg[2]; ++n2) {
                double tv[3];
                image_vector(cr, n0, n1, n2, tv);
                const double rv[3] = {x[3 * a] - x[3 * o] - tv[0], x[3 * a + 1] - x[3 * o + 1] - tv[1],
                                      x[3 * a + 2] - x[3 * o + 2] - tv[2]};
                const double r2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
                if (r2 <= kPairEps2 || r2 > c2) continue;
                const double r = sqrt(r2);
                if (r <= reps) continue;
                double v, dv;
                count_fn(r, rc, v, dv);
                const double pref = dedcn_ao * dv / r;
                for (int d = 0; d < 3; ++d) acc[d] += rv[d] * pref;
            }
    }
    for (int d = 0; d < 3; ++d) acc[d] = wave_sum(acc[d]);
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) g[3 * a + d] += acc[d];
}

// Halogen bond of a periodic instance (halogen.rs:64-160 over the triples of
// :716-740): the donor's neighbours are every (atom, image) within 20 bohr but
// the donor itself in the origin cell; the nearest one that is not an image of
// the donor is its bonded neighbour; acceptors may be images.  Forces go to the
// atoms the images belong to.
__device__ void halogen_periodic(const DeviceBatch& b, const DeviceTables& t, int inst, double* e_hal,
                                 double* g_hal) {
    const int a0 = b.atom0[inst], a1 = b.atom0[inst + 1];
    const double* x = b.atom_x;
    const double damp = t.glob.xbdamp, rads = t.glob.xbrad;
    const double c2 = 20.0 * 20.0, EPS = 1e-18, ALP = 6.0, LJ2 = 6.0;
    const CellRec& cr = b.cell[inst];
    int rg[3];
    image_ranges(cr, 20.0, rg);
    double e = 0.0;
    for (int X = a0; X < a1; ++X) {
        const int zx = b.atom_z[X];
        if (!(zx == 17 || zx == 35 || zx == 53 || zx == 85)) continue;
        int K = -1;
        double best = INFINITY, tk[3] = {0, 0, 0};
        for (int n0 = -rg[0]; n0 <= rg[0]; ++n0)
            for (int n1 = -rg[1]; n1 <= rg[1]; ++n1)
                for (int n2 = -rg[2]; n2 <= rg[2]; ++n2) {
                    double tv[3];
                    image_vector(cr, n0, n1, n2, tv);
                    for (int j = a0; j < a1; ++j) {
                        if (j == X) continue;
                        double d[3];
                        for (int q = 0; q < 3; ++q) d[q] = x[3 * j + q] + tv[q] - x[3 * X + q];
                        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                        if (r2 <= kPairEps2 || r2 > c2) continue;
                        if (r2 < best) {
                            best = r2;
                            K = j;
                            for (int q = 0; q < 3; ++q) tk[q] = tv[q];
                        }
                    }
                }
        if (K < 0) continue;
        const double cc = t.elements[zx].xb_strength;
        if (cc == 0.0) continue;
        for (int n0 = -rg[0]; n0 <= rg[0]; ++n0)
            for (int n1 = -rg[1]; n1 <= rg[1]; ++n1)
                for (int n2 = -rg[2]; n2 <= rg[2]; ++n2) {
                    double tv[3];
                    image_vector(cr, n0, n1, n2, tv);
                    for (int j = a0; j < a1; ++j) {
                        const int zj = b.atom_z[j];
                        if (!(zj == 7 || zj == 8 || zj == 15 || zj == 16)) continue;
                        double dxj[3], dxk[3], dkj[3];
                        for (int d = 0; d < 3; ++d) {
                            dxj[d] = x[3 * j + d] + tv[d] - x[3 * X + d];
                            dxk[d] = x[3 * K + d] + tk[d] - x[3 * X + d];
                            dkj[d] = x[3 * j + d] + tv[d] - x[3 * K + d] - tk[d];
                        }
                        const double d2jx = dxj[0] * dxj[0] + dxj[1] * dxj[1] + dxj[2] * dxj[2];
                        if (d2jx <= kPairEps2 || d2jx > c2) continue;
                        const double d2kx = dxk[0] * dxk[0] + dxk[1] * dxk[1] + dxk[2] * dxk[2];
                        const double d2jk = dkj[0] * dkj[0] + dkj[1] * dkj[1] + dkj[2] * dkj[2];
                        if (d2jx <= EPS || d2kx <= EPS) continue;
                        const double r0 = rads * (t.elements[zx].atomic_radius + t.elements[zj].atomic_radius);
                        const double rjx = sqrt(d2jx) + EPS, rkx = sqrt(d2kx) + EPS;
                        const double xy = sqrt(d2kx * d2jx);
                        if (xy <= EPS) continue;
                        const double term = (d2kx + d2jx - d2jk) / xy;
                        const double ab = 0.5 - 0.25 * term;
                        const double aterm = pow(ab, ALP);
                        const double t13 = r0 / rjx;
                        const double t14lj = pow(t13, 12.0);
                        e += aterm * cc * (t14lj - damp * pow(t13, LJ2)) / (1.0 + t14lj);
                        const double t14 = pow(r0 / rjx, LJ2);
                        const double num = t14 * t14 - damp * t14, den = 1.0 + t14 * t14;
                        const double tlj = num / den;
                        double dt = 2.0 * LJ2 * num * t14 * t14 / (rjx * den * den);
                        dt += LJ2 * t14 * (damp - 2.0 * t14) / (rjx * den);
                        dt *= aterm * cc / rjx;
                        const double pf = -0.25 * ALP * pow(ab, ALP - 1.0) * cc * tlj;
                        const double djx = (2.0 / rkx - term / rjx) * pf / rjx;
                        const double dkx = (2.0 / rjx - term / rkx) * pf / rkx;
                        const double djk = 2.0 * pf / xy;
                        for (int d = 0; d < 3; ++d) {
                            g_hal[3 * j + d] += dxj[d] * dt + dxj[d] * djx - dkj[d] * djk;
                            g_hal[3 * X + d] -= dxj[d] * dt + dxj[d] * djx + dxk[d] * dkx;
                            g_hal[3 * K + d] += dxk[d] * dkx + dkj[d] * djk;
                        }
                    }
                }
    }
    e_hal[inst] = e;
}

// Halogen-bond correction (halogen.rs:64-160, :716-784); one thread per instance.
__global__ void k_halogen(DeviceBatch b, DeviceTables t, double* e_hal, double* g_hal) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= b.ninst) return;
    if (periodic_instance(b, inst)) {
        halogen_periodic(b, t, inst, e_hal, g_hal);
        return;
    }
    const int a0 = b.atom0[inst], a1 = b.atom0[inst + 1];
    const double* x = b.atom_x;
    const double damp = t.glob.xbdamp, rads = t.glob.xbrad;
    const double c2 = 20.0 * 20.0, EPS = 1e-18, ALP = 6.0, LJ2 = 6.0;
    double e = 0.0;
    for (int X = a0; X < a1; ++X) {
        const int zx = b.atom_z[X];
        if (!(zx == 17 || zx == 35 || zx == 53 || zx == 85)) continue;
        int K = -1;
        double best = INFINITY;
        for (int j = a0; j < a1; ++j) {
            if (j == X) continue;
            const double r2 = dist2(x, X, j);
            if (r2 <= kPairEps2 || r2 > c2) continue;
            if (r2 < best) { best = r2; K = j; }
        }
        if (K < 0) continue;
        const double cc = t.elements[zx].xb_strength;
        if (cc == 0.0) continue;
        for (int j = a0; j < a1; ++j) {
            if (j == X) continue;
            const double r2 = dist2(x, X, j);
            if (r2 <= kPairEps2 || r2 > c2) continue;
            const int zj = b.atom_z[j];
            if (!(zj == 7 || zj == 8 || zj == 15 || zj == 16)) continue;
            const double r0 = rads * (t.elements[zx].atomic_radius + t.elements[zj].atomic_radius);
            double dxj[3], dxk[3], dkj[3];
            for (int d = 0; d < 3; ++d) {
                dxj[d] = x[3 * j + d] - x[3 * X + d];
                dxk[d] = x[3 * K + d] - x[3 * X + d];
                dkj[d] = x[3 * j + d] - x[3 * K + d];
            }
            const double d2jx = dxj[0] * dxj[0] + dxj[1] * dxj[1] + dxj[2] * dxj[2];
            const double d2kx = dxk[0] * dxk[0] + dxk[1] * dxk[1] + dxk[2] * dxk[2];
            const double d2jk = dkj[0] * dkj[0] + dkj[1] * dkj[1] + dkj[2] * dkj[2];
            if (d2jx <= EPS || d2kx <= EPS) continue;
            const double rjx = sqrt(d2jx) + EPS, rkx = sqrt(d2kx) + EPS;
            const double xy = sqrt(d2kx * d2jx);
            if (xy <= EPS) continue;
            const double term = (d2kx + d2jx - d2jk) / xy;
            const double ab = 0.5 - 0.25 * term;
            const double aterm = pow(ab, ALP);
            const double t13 = r0 / rjx;
            const double t14lj = pow(t13, 12.0);
            e += aterm * cc * (t14lj - damp * pow(t13, LJ2)) / (1.0 + t14lj);
            const double t14 = pow(r0 / rjx, LJ2);
            const double num = t14 * t14 - damp * t14, den = 1.0 + t14 * t14;
            const double tlj = num / den;
            double dt = 2.0 * LJ2 * num * t14 * t14 / (rjx * den * den);
            dt += LJ2 * t14 * (damp - 2.0 * t14) / (rjx * den);
            dt *= aterm * cc / rjx;
            const double pf = -0.25 * ALP * pow(ab, ALP - 1.0) * cc * tlj;
            const double djx = (2.0 / rkx - term / rjx) * pf / rjx;
            const double dkx = (2.0 / rjx - term / rkx) * pf / rkx;
            const double djk = 2.0 * pf / xy;
            for (int d = 0; d < 3; ++d) {
                g_hal[3 * j + d] += dxj[d] * dt + dxj[d] * djx - dkj[d] * djk;
                g_hal[3 * X + d] -= dxj[d] * dt + dxj[d] * djx + dxk[d] * dkx;
                g_hal[3 * K + d] += dxk[d] * dkx + dkj[d] * djk;
            }
        }
    }
    e_hal[inst] = e;
}

// Per-instance sums of per-atom values (fixed order).
__global__ void k_instance_sum(DeviceBatch b, const double* per_atom, double* out) {
    __shared__ double lds[16];
    const int inst = blockIdx.x;
    const int a0 = b.atom0[inst], a1 = b.atom0[inst + 1];
    double s = 0.0;
    for (int a = a0 + threadIdx.x; a < a1; a += blockDim.x) s += per_atom[a];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) out[inst] = s;
}

// Electronic gradient, per-atom part: diagonal dE/dCN (gradient.rs:465-469) and
// the isotropic Coulomb term (gradient.rs:493-516).
__global__ void k_grad_atoms(DeviceBatch b, DeviceTables t, const double* q, int cn_on, int scc_on,
                             double* dedcn_h, double* g_el, int natoms) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms) return;
    const int inst = b.atom_inst[a];
    const int ld = b.ld[inst];
    const int za = b.atom_z[a];
    const int nsa = t.elements[za].nshell;
    const int sa0 = b.atom_sh0[a];
    gptr<const double> P = (gptr<const double>)b.P[inst];
    if (cn_on) {
        double d = dedcn_h[a];
        for (int k = 0; k < nsa; ++k) {
            const int sa = sa0 + k;
            const ShellType& A = t.shells[b.sh_type[sa]];
            for (int m = 0; m < A.nao; ++m) {
                const int mu = b.sh_ao[sa] + m;
                d += (-A.kcn) * P[(size_t)mu * ld + mu];
            }
        }
        dedcn_h[a] = d;
    }
    // periodic instances: the Ewald gradient kernel (periodic.hip) instead
    if (!scc_on || periodic_instance(b, inst)) return;
    const int s0 = b.sh0[inst], s1 = s0 + b.nsh[inst];
    const double* x = b.atom_x;
    double acc[3] = {0, 0, 0};
    for (int k = 0; k < nsa; ++k) {
        const int sa = sa0 + k;
        const double ha = t.shells[b.sh_type[sa]].hardness;
        for (int sb = s0; sb < s1; ++sb) {
            const int o = b.sh_atom[sb];
            if (o == a) continue;
            double rv[3];
            for (int d = 0; d < 3; ++d) rv[d] = x[3 * a + d] - x[3 * o + d];
            const double r = sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
            if (r <= 1e-12) continue;
            const double hb = t.shells[b.sh_type[sb]].hardness;
            const double h = 2.0 / (1.0 / ha + 1.0 / hb);
            const double den = r * r + 1.0 / (h * h);
            const double pref = -1.0 / (den * sqrt(den)) * (q[sa] * q[sb]);
            for (int d = 0; d < 3; ++d) acc[d] += rv[d] * pref;
        }
    }
    for (int d = 0; d < 3; ++d) g_el[3 * a + d] += acc[d];
}


// ---------------------------------------------------------------------------
// Shell-pair integrals specialised by angular-momentum class (LA, LB).
// Overlap block <A mu|B nu> (integrals.rs:290-360) and, with DERIV, its
// derivative with respect to the bra centre (integrals.rs:481-537, :782-868),
// with the primitive screening bound of integrals.rs:30-45.
// Block layout: S[mu * NB + nu], dA[(mu * NB + nu) * 3 + d].
namespace {

template <int LA, int LB, bool DERIV>
__device__ __forceinline__ void shell_block(const ShellType& A, const ShellType& B, const double* ra,
                                            const double* rb, double* __restrict__ S, double* __restrict__ dA) {
    constexpr int NA = 2 * LA + 1, NB = 2 * LB + 1;
    constexpr int IM = LA + (DERIV ? 1 : 0);
    constexpr int GM = (IM + LB) / 2 + 1;
    constexpr int order = LA + LB + (DERIV ? 3 : 2);
#pragma unroll
    for (int k = 0; k < NA * NB; ++k) S[k] = 0.0;
    if constexpr (DERIV) {
#pragma unroll
        for (int k = 0; k < 3 * NA * NB; ++k) dA[k] = 0.0;
    }
    const double dx = ra[0] - rb[0], dy = ra[1] - rb[1], dz = ra[2] - rb[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    const double csmax = ao_abs_sum(LA, 1) * ao_abs_sum(LB, 1);
    for (int ip = 0; ip < A.nprim; ++ip) {
        const double al = A.alpha[ip], cai = A.coeff[ip];
        for (int jp = 0; jp < B.nprim; ++jp) {
            const double be = B.alpha[jp];
            const double pref = cai * B.coeff[jp];
            // one division and one square root per primitive pair: 1/p and sqrt(1/p)
            const double p = al + be, ip = 1.0 / p;
            const double kab = exp(-al * be * ip * r2);
            const double spread = sqrt(ip);
            const double g0 = kSqrtPiK * spread;  // sqrt(pi / p)
            const double gpref = kPi * ip * g0;   // (pi / p)^(3/2)
            const double poly = ipow(1.0 + r + spread, order);
            if (fabs(pref) * csmax * kab * gpref * poly < kPrimScreen) continue;
            double g[GM];
            g[0] = g0;
#pragma unroll
            for (int m = 1; m < GM; ++m) g[m] = g[m - 1] * (m - 0.5) * ip;
            double mt[3][IM + 1][LB + 1];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double P = (al * ra[d] + be * rb[d]) * ip;
                const double pa = P - ra[d], pb = P - rb[d];
#pragma unroll
                for (int i = 0; i <= IM; ++i)
#pragma unroll
                    for (int j = 0; j <= LB; ++j) mt[d][i][j] = moment1d(i, j, pa, pb, g);
            }
#pragma unroll
            for (int mu = 0; mu < NA; ++mu) {
                CartTerm ta[3];
                const int nta = ao_terms(LA, mu, ta);
#pragma unroll
                for (int nu = 0; nu < NB; ++nu) {
                    if constexpr (LA == 2 || LB == 2) {
                        const double cs = ao_abs_sum(LA, mu) * ao_abs_sum(LB, nu);
                        if (fabs(pref) * cs * kab * gpref * poly < kPrimScreen) continue;
                    }
                    CartTerm tb[3];
                    const int ntb = ao_terms(LB, nu, tb);
                    double sv = 0.0, ga0 = 0.0, ga1 = 0.0, ga2 = 0.0;
                    for (int a = 0; a < nta; ++a)
                        for (int bq = 0; bq < ntb; ++bq) {
                            const double cf = pref * ta[a].c * tb[bq].c * kab;
                            const int i0 = ta[a].px, i1 = ta[a].py, i2 = ta[a].pz;
                            const int j0 = tb[bq].px, j1 = tb[bq].py, j2 = tb[bq].pz;
                            const double m0 = mt[0][i0][j0], m1 = mt[1][i1][j1], m2 = mt[2][i2][j2];
                            sv += cf * (m0 * m1 * m2);
                            if constexpr (DERIV) {
                                const double d0 = 2.0 * al * mt[0][i0 + 1][j0] - (i0 > 0 ? i0 * mt[0][i0 > 0 ? i0 - 1 : 0][j0] : 0.0);
                                const double d1 = 2.0 * al * mt[1][i1 + 1][j1] - (i1 > 0 ? i1 * mt[1][i1 > 0 ? i1 - 1 : 0][j1] : 0.0);
                                const double d2 = 2.0 * al * mt[2][i2 + 1][j2] - (i2 > 0 ? i2 * mt[2][i2 > 0 ? i2 - 1 : 0][j2] : 0.0);
                                ga0 += cf * (d0 * (m1 * m2));
                                ga1 += cf * (d1 * (m2 * m0));
                                ga2 += cf * (d2 * (m0 * m1));
                            }
                        }
                    S[mu * NB + nu] += sv;
                    if constexpr (DERIV) {
                        dA[(mu * NB + nu) * 3 + 0] += ga0;
                        dA[(mu * NB + nu) * 3 + 1] += ga1;
                        dA[(mu * NB + nu) * 3 + 2] += ga2;
                    }
                }
            }
        }
    }
}

// S and H0 for one (shell of A, shell of B) pair with B <= A; writes both triangles.
template <int LA, int LB>
__device__ __forceinline__ void s_h0_pair(const DeviceBatch& b, const DeviceTables& t, const double* cn, double c2,
                                          int a, int sa, int sb, int ld, gptr<double> S, gptr<double> H) {
    const int ob = b.sh_atom[sb];
    const ShellType& A = t.shells[b.sh_type[sa]];
    const ShellType& B = t.shells[b.sh_type[sb]];
    const double* x = b.atom_x;
    const double ra[3] = {x[3 * a], x[3 * a + 1], x[3 * a + 2]};
    const double rb[3] = {x[3 * ob], x[3 * ob + 1], x[3 * ob + 2]};
    constexpr int NA = 2 * LA + 1, NB = 2 * LB + 1;
    double blk[NA * NB];
    double r2 = 0.0;
    bool inside = true;
    if (ob != a) {
        r2 = dist2(x, a, ob);
        inside = r2 > kPairEps2 && r2 <= c2;
    }
    if (inside) shell_block<LA, LB, false>(A, B, ra, rb, blk, nullptr);
    else
        for (int k = 0; k < NA * NB; ++k) blk[k] = 0.0;
    const int za = b.atom_z[a], zb = b.atom_z[ob];
    const double sea = A.hdiag - A.kcn * cn[a];
    const double seb = B.hdiag - B.kcn * cn[ob];
    double h;
    if (ob == a) {
        h = 0.5 * (sea + seb);
    } else {
        const double rr = sqrt(sqrt(r2) / (t.elements[za].atomic_radius + t.elements[zb].atomic_radius));
        h = 0.5 * (sea + seb) * hscale(t, za, A, zb, B) * ((1.0 + A.poly * rr) * (1.0 + B.poly * rr));
    }
    const int mu0 = b.sh_ao[sa], nu0 = b.sh_ao[sb];
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const double sv = blk[i * NB + j];
            const size_t e1 = (size_t)(mu0 + i) * ld + (nu0 + j);  // element (nu, mu)
            const size_t e2 = (size_t)(nu0 + j) * ld + (mu0 + i);  // element (mu, nu)
            S[e1] = sv;
            S[e2] = sv;
            H[e1] = sv * h;
            H[e2] = sv * h;
        }
}

// One image of the (shell sa of A, shell sb of B) pair of a periodic instance
// (pbc/bloch.rs:58-236): the overlap block of B at x_B + T and the H0 block,
// the on-site formula for B = A in the origin cell (same), the off-site one at
// the image distance otherwise.  bs[i NB + j], bh[i NB + j].
template <int LA, int LB>
__device__ __forceinline__ void s_h0_image(const DeviceBatch& b, const DeviceTables& t, const double* cn,
                                           const CellRec& cr, int a, int sa, int sb, int n0, int n1, int n2,
                                           double* bs, double* bh) {
    const int ob = b.sh_atom[sb];
    const ShellType& A = t.shells[b.sh_type[sa]];
    const ShellType& B = t.shells[b.sh_type[sb]];
    const double* x = b.atom_x;
    const double ra[3] = {x[3 * a], x[3 * a + 1], x[3 * a + 2]};
    double tv[3];
    image_vector(cr, n0, n1, n2, tv);
    const double rb[3] = {x[3 * ob] + tv[0], x[3 * ob + 1] + tv[1], x[3 * ob + 2] + tv[2]};
    const bool same = ob == a && n0 == 0 && n1 == 0 && n2 == 0;
    shell_block<LA, LB, false>(A, B, ra, rb, bs, nullptr);
    const int za = b.atom_z[a], zb = b.atom_z[ob];
    double h = 0.5 * ((A.hdiag - A.kcn * cn[a]) + (B.hdiag - B.kcn * cn[ob]));
    if (!same) {
        const double dx = ra[0] - rb[0], dy = ra[1] - rb[1], dz = ra[2] - rb[2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double rr = sqrt(sqrt(r2) / (t.elements[za].atomic_radius + t.elements[zb].atomic_radius));
        h *= hscale(t, za, A, zb, B) * ((1.0 + A.poly * rr) * (1.0 + B.poly * rr));
    }
    constexpr int NA = 2 * LA + 1, NB = 2 * LB + 1;
#pragma unroll
    for (int k = 0; k < NA * NB; ++k) bh[k] = bs[k] * h;
}

// Sums K values over the runs of equal keys of a wavefront (keys non-decreasing
// over the lanes, -1 for none): a segmented scan in a fixed order, the run's sum
// ending on its last lane, for which the return value is true.
template <int K>
__device__ __forceinline__ bool run_sums(int key, double* v) {
    const int lane = threadIdx.x & 63;
    for (int off = 1; off < 64; off <<= 1) {
        const int ku = __shfl_up(key, off, 64);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double vu = __shfl_up(v[k], off, 64);
            if (lane >= off && ku == key) v[k] += vu;
        }
    }
    const int kn = __shfl_down(key, 1, 64);
    return key >= 0 && (lane == 63 || kn != key);
}

// S(Gamma) and H0(Gamma) of shell sa of atom a against the shells of class LB
// with index <= sa (sh_by_l[s_first .. s_first + nvalid)), over the images of
// the 30-bohr cutoff outside the overlap screen (pbc/bloch.rs:58-236), both
// triangles.  The (shell, image) candidates are tested 64 at a time, the inside
// ones compacted into the LDS queue pq in candidate order and evaluated one per
// lane; the lanes' blocks are summed per shell by run_sums and added to the
// blocks, which are zeroed first.
template <int LA, int LB>
__device__ void s_h0_shell_pbc(const DeviceBatch& b, const DeviceTables& t, const double* cn, const CellRec& cr,
                               const int* rg, int a, int sa, int s_first, int nvalid, int ld, gptr<double> S,
                               gptr<double> H, int* pq) {
    constexpr int NA = 2 * LA + 1, NB = 2 * LB + 1;
    const int lane = threadIdx.x;
    const int mu0 = b.sh_ao[sa];
    // zero the blocks (sa, sb) of both triangles
    for (int q = lane; q < nvalid * NA * NB; q += 64) {
        const int sb = b.sh_by_l[s_first + q / (NA * NB)];
        const int i = (q % (NA * NB)) / NB, j = q % NB;
        const int nu0 = b.sh_ao[sb];
        S[(size_t)(mu0 + i) * ld + (nu0 + j)] = 0.0;
        H[(size_t)(mu0 + i) * ld + (nu0 + j)] = 0.0;
        S[(size_t)(nu0 + j) * ld + (mu0 + i)] = 0.0;
        H[(size_t)(nu0 + j) * ld + (mu0 + i)] = 0.0;
    }
    __syncthreads();
    const int d1 = 2 * rg[1] + 1, d2 = 2 * rg[2] + 1, nimg = (2 * rg[0] + 1) * d1 * d2;
    const int ncand = nvalid * nimg;
    const double* x = b.atom_x;
    const int za = b.atom_z[a];
    const double c2 = kPbcAoCutoff * kPbcAoCutoff;
    auto evaluate_and_add = [&](bool have, int c) {
        double v[2 * NA * NB];
        int key = -1, sb = 0;
        if (have) {
            sb = b.sh_by_l[s_first + c / nimg];
            const int im = c % nimg;
            s_h0_image<LA, LB>(b, t, cn, cr, a, sa, sb, im / (d1 * d2) - rg[0], (im / d2) % d1 - rg[1],
                               im % d2 - rg[2], v, v + NA * NB);
            key = c / nimg;
        } else {
#pragma unroll
            for (int k = 0; k < 2 * NA * NB; ++k) v[k] = 0.0;
        }
        if (run_sums<2 * NA * NB>(key, v)) {
            const int nu0 = b.sh_ao[sb];
#pragma unroll
            for (int i = 0; i < NA; ++i)
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    const size_t e1 = (size_t)(mu0 + i) * ld + (nu0 + j);
                    S[e1] += v[i * NB + j];
                    H[e1] += v[NA * NB + i * NB + j];
                    if (sb != sa) {  // the diagonal block's mirror is itself
                        const size_t e2 = (size_t)(nu0 + j) * ld + (mu0 + i);
                        S[e2] += v[i * NB + j];
                        H[e2] += v[NA * NB + i * NB + j];
                    }
                }
        }
    };
    int cursor = 0;  // queued candidates, uniform over the wavefront
    for (int c0 = 0; c0 < ncand; c0 += 64) {
        const int c = c0 + lane;
        bool inside = false;
        if (c < ncand) {
            const int sb = b.sh_by_l[s_first + c / nimg];
            const int ob = b.sh_atom[sb];
            const int im = c % nimg;
            const int n0 = im / (d1 * d2) - rg[0], n1 = (im / d2) % d1 - rg[1], n2 = im % d2 - rg[2];
            double tv[3];
            image_vector(cr, n0, n1, n2, tv);
            const double dx = x[3 * a] - (x[3 * ob] + tv[0]), dy = x[3 * a + 1] - (x[3 * ob + 1] + tv[1]),
                         dz = x[3 * a + 2] - (x[3 * ob + 2] + tv[2]);
            const double r2 = dx * dx + dy * dy + dz * dz;
            const bool same = ob == a && n0 == 0 && n1 == 0 && n2 == 0;
            inside = same || (r2 > kPairEps2 && r2 <= c2 && !overlap_screened(t, za, b.atom_z[ob], r2));
        }
        const unsigned long long m = __ballot(inside);
        if (inside) pq[cursor + __popcll(m & ((1ull << lane) - 1ull))] = c;
        cursor += __popcll(m);
        __syncthreads();
        if (cursor >= 64) {
            evaluate_and_add(true, pq[lane]);
            __syncthreads();
            if (lane < cursor - 64) pq[lane] = pq[64 + lane];
            cursor -= 64;
            __syncthreads();
        }
    }
    evaluate_and_add(lane < cursor, lane < cursor ? pq[lane] : 0);
    __syncthreads();
}

// Electronic gradient contribution of one (shell of A, shell of B) pair to atom A
// (gradient.rs:358-455): Pulay/shell-potential term, polynomial term, dE/dCN_A.
// The AO-pair weights oc = P (2 h - v_mu - v_nu) - 2 W and P are applied inside
// the primitive loop, so neither the overlap block nor its derivative block
// is ever stored.
// Per-atom gradient accumulator: plain FP64 sums, or FP32 sums with Kahan
// compensation (mixed precision).
template <typename R>
struct GradAcc {
    R g[3] = {0, 0, 0}, dc = 0;
    R cg[3] = {0, 0, 0}, cdc = 0;
    __device__ __forceinline__ static void add(R& s, R& c, R v) {
        if constexpr (sizeof(R) == sizeof(float)) {
            const R y = v - c;
            const R t = s + y;
            c = (t - s) - y;
            s = t;
        } else {
            s += v;
        }
    }
    __device__ __forceinline__ double value(int d) const { return (double)g[d] - (double)cg[d]; }
    __device__ __forceinline__ double dcn() const { return (double)dc - (double)cdc; }
};

// PBC: B sits at its image x_B + tv (P(T) = P at the Gamma point); for an
// image of A itself (self) only dE/dCN_A is accumulated: the pair's distance
// does not move with A.
template <int LA, int LB, typename R, bool PBC = false>
__device__ __forceinline__ void grad_pair(const DeviceBatch& b, const DeviceTables& t, const double* cn,
                                          const double* vsh, int scc_on, int cn_on, int a, int sa, int sb, int ld,
                                          gptr<const double> P, gptr<const double> W, GradAcc<R>& acc,
                                          const double* tv = nullptr, bool self = false, double* oth = nullptr) {
    const int ob = b.sh_atom[sb];
    const ShellType& A = t.shells[b.sh_type[sa]];
    const ShellType& B = t.shells[b.sh_type[sb]];
    const double* x = b.atom_x;
    const double ra[3] = {x[3 * a], x[3 * a + 1], x[3 * a + 2]};
    double rb[3] = {x[3 * ob], x[3 * ob + 1], x[3 * ob + 2]};
    if constexpr (PBC)
        for (int k = 0; k < 3; ++k) rb[k] += tv[k];
    constexpr int NA = 2 * LA + 1, NB = 2 * LB + 1;
    const int za = b.atom_z[a], zb = b.atom_z[ob];
    const double dx = ra[0] - rb[0], dy = ra[1] - rb[1], dz = ra[2] - rb[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    const double rr = sqrt(r / (t.elements[za].atomic_radius + t.elements[zb].atomic_radius));
    const double fa = 1.0 + A.poly * rr, fb = 1.0 + B.poly * rr;
    const double hs = hscale(t, za, A, zb, B) * (fa * fb);
    const double dse_a = cn_on ? -A.kcn : 0.0, dse_b = cn_on ? -B.kcn : 0.0;
    const double hij = 0.5 * ((A.hdiag + dse_a * cn[a]) + (B.hdiag + dse_b * cn[ob])) * hs;
    const double vv = scc_on ? vsh[sa] + vsh[sb] : 0.0;
    const int mu0 = b.sh_ao[sa], nu0 = b.sh_ao[sb];
    // contraction coefficients P (2 h - v_mu - v_nu) - 2 W (gradient.rs:400-440)
    R oc[NA * NB], pw[NA * NB];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        gptr<const double> Pc = P + (size_t)(mu0 + i) * ld + nu0;
        gptr<const double> Wc = W + (size_t)(mu0 + i) * ld + nu0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const double p = Pc[j];
            pw[i * NB + j] = (R)p;
            oc[i * NB + j] = (R)(p * (2.0 * hij - vv) - 2.0 * Wc[j]);
        }
    }
    constexpr int IM = LA + 1;
    constexpr int GM = (IM + LB) / 2 + 1;
    constexpr int order = LA + LB + 3;
    const double csmax = ao_abs_sum(LA, 1) * ao_abs_sum(LB, 1);
    R g0 = 0, g1 = 0, g2 = 0, sp = 0;
    if constexpr (sizeof(R) == sizeof(float)) {
        // Mixed precision: the primitive integrals, their derivatives and the
        // contraction all in FP32.  The product-centre offsets come from the
        // bond vector (P - A = be (B - A) / p), not from absolute positions,
        // so no cancellation is lost in FP32.
        const float r2f = (float)r2, rf = (float)r;
        const float bab[3] = {(float)(rb[0] - ra[0]), (float)(rb[1] - ra[1]), (float)(rb[2] - ra[2])};
        const float csf = (float)csmax;
        for (int ip = 0; ip < A.nprim; ++ip) {
            const float al = (float)A.alpha[ip], cai = (float)A.coeff[ip];
            for (int jp = 0; jp < B.nprim; ++jp) {
                const float be = (float)B.alpha[jp];
                const float pref = cai * (float)B.coeff[jp];
                const float p = al + be, ip_ = 1.0f / p;
                const float kab = expf(-al * be * ip_ * r2f);
                const float spread = sqrtf(ip_);
                const float gpref = ((float)kPi * ip_) * sqrtf((float)kPi * ip_);
                float poly = 1.0f;
                for (int q = 0; q < order; ++q) poly *= 1.0f + rf + spread;
                if (fabsf(pref) * csf * kab * gpref * poly < (float)kPrimScreen) continue;
                float gm[GM];
                gm[0] = sqrtf((float)kPi * ip_);
#pragma unroll
                for (int m = 1; m < GM; ++m) gm[m] = gm[m - 1] * ((float)m - 0.5f) * ip_;
                float mt[3][IM + 1][LB + 1];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const float pa = be * ip_ * bab[d], pb = -al * ip_ * bab[d];
#pragma unroll
                    for (int i = 0; i <= IM; ++i)
#pragma unroll
                        for (int j = 0; j <= LB; ++j) mt[d][i][j] = moment1d_t<float>(i, j, pa, pb, gm);
                }
#pragma unroll
                for (int mu = 0; mu < NA; ++mu) {
                    CartTerm ta[3];
                    const int nta = ao_terms(LA, mu, ta);
#pragma unroll
                    for (int nu = 0; nu < NB; ++nu) {
                        if constexpr (LA == 2 || LB == 2) {
                            const float cs = (float)(ao_abs_sum(LA, mu) * ao_abs_sum(LB, nu));
                            if (fabsf(pref) * cs * kab * gpref * poly < (float)kPrimScreen) continue;
                        }
                        CartTerm tb[3];
                        const int ntb = ao_terms(LB, nu, tb);
                        float sv = 0.0f, ga0 = 0.0f, ga1 = 0.0f, ga2 = 0.0f;
                        for (int aq = 0; aq < nta; ++aq)
                            for (int bq = 0; bq < ntb; ++bq) {
                                const float cf = pref * (float)(ta[aq].c * tb[bq].c) * kab;
                                const int i0 = ta[aq].px, i1 = ta[aq].py, i2 = ta[aq].pz;
                                const int j0 = tb[bq].px, j1 = tb[bq].py, j2 = tb[bq].pz;
                                const float m0 = mt[0][i0][j0], m1 = mt[1][i1][j1], m2 = mt[2][i2][j2];
                                sv += cf * (m0 * m1 * m2);
                                const float d0 = 2.0f * al * mt[0][i0 + 1][j0] - (i0 > 0 ? (float)i0 * mt[0][i0 > 0 ? i0 - 1 : 0][j0] : 0.0f);
                                const float d1 = 2.0f * al * mt[1][i1 + 1][j1] - (i1 > 0 ? (float)i1 * mt[1][i1 > 0 ? i1 - 1 : 0][j1] : 0.0f);
                                const float d2 = 2.0f * al * mt[2][i2 + 1][j2] - (i2 > 0 ? (float)i2 * mt[2][i2 > 0 ? i2 - 1 : 0][j2] : 0.0f);
                                ga0 += cf * (d0 * (m1 * m2));
                                ga1 += cf * (d1 * (m2 * m0));
                                ga2 += cf * (d2 * (m0 * m1));
                            }
                        const R w = oc[mu * NB + nu];
                        g0 += ga0 * w;
                        g1 += ga1 * w;
                        g2 += ga2 * w;
                        sp += sv * pw[mu * NB + nu];
                    }
                }
            }
        }
    } else {
        for (int ip = 0; ip < A.nprim; ++ip) {
            const double al = A.alpha[ip], cai = A.coeff[ip];
            for (int jp = 0; jp < B.nprim; ++jp) {
                const double be = B.alpha[jp];
                const double pref = cai * B.coeff[jp];
                // one division and one square root per primitive pair: 1/p and sqrt(1/p)
                const double p = al + be, ip = 1.0 / p;
                const double kab = exp(-al * be * ip * r2);
                const double spread = sqrt(ip);
                const double sqp = kSqrtPiK * spread;  // sqrt(pi / p)
                const double gpref = kPi * ip * sqp;   // (pi / p)^(3/2)
                const double poly = ipow(1.0 + r + spread, order);
                if (fabs(pref) * csmax * kab * gpref * poly < kPrimScreen) continue;
                double gm[GM];
                gm[0] = sqp;
    #pragma unroll
                for (int m = 1; m < GM; ++m) gm[m] = gm[m - 1] * (m - 0.5) * ip;
                double mt[3][IM + 1][LB + 1];
    #pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double Pc = (al * ra[d] + be * rb[d]) * ip;
                    const double pa = Pc - ra[d], pb = Pc - rb[d];
    #pragma unroll
                    for (int i = 0; i <= IM; ++i)
    #pragma unroll
                        for (int j = 0; j <= LB; ++j) mt[d][i][j] = moment1d(i, j, pa, pb, gm);
                }
    #pragma unroll
                for (int mu = 0; mu < NA; ++mu) {
                    CartTerm ta[3];
                    const int nta = ao_terms(LA, mu, ta);
    #pragma unroll
                    for (int nu = 0; nu < NB; ++nu) {
                        if constexpr (LA == 2 || LB == 2) {
                            const double cs = ao_abs_sum(LA, mu) * ao_abs_sum(LB, nu);
                            if (fabs(pref) * cs * kab * gpref * poly < kPrimScreen) continue;
                        }
                        CartTerm tb[3];
                        const int ntb = ao_terms(LB, nu, tb);
                        double sv = 0.0, ga0 = 0.0, ga1 = 0.0, ga2 = 0.0;
                        for (int aq = 0; aq < nta; ++aq)
                            for (int bq = 0; bq < ntb; ++bq) {
                                const double cf = pref * ta[aq].c * tb[bq].c * kab;
                                const int i0 = ta[aq].px, i1 = ta[aq].py, i2 = ta[aq].pz;
                                const int j0 = tb[bq].px, j1 = tb[bq].py, j2 = tb[bq].pz;
                                const double m0 = mt[0][i0][j0], m1 = mt[1][i1][j1], m2 = mt[2][i2][j2];
                                sv += cf * (m0 * m1 * m2);
                                const double d0 = 2.0 * al * mt[0][i0 + 1][j0] - (i0 > 0 ? i0 * mt[0][i0 > 0 ? i0 - 1 : 0][j0] : 0.0);
                                const double d1 = 2.0 * al * mt[1][i1 + 1][j1] - (i1 > 0 ? i1 * mt[1][i1 > 0 ? i1 - 1 : 0][j1] : 0.0);
                                const double d2 = 2.0 * al * mt[2][i2 + 1][j2] - (i2 > 0 ? i2 * mt[2][i2 > 0 ? i2 - 1 : 0][j2] : 0.0);
                                ga0 += cf * (d0 * (m1 * m2));
                                ga1 += cf * (d1 * (m2 * m0));
                                ga2 += cf * (d2 * (m0 * m1));
                            }
                        const R w = oc[mu * NB + nu];
                        g0 += (R)ga0 * w;
                        g1 += (R)ga1 * w;
                        g2 += (R)ga2 * w;
                        sp += (R)sv * pw[mu * NB + nu];
                    }
                }
            }
        }
    }
    // polynomial term 2 P hij S dln(poly)/dR_A (gradient.rs:1157-1175) and dE/dCN_A
    const double pl = fa * fb;
    double dlp0 = 0.0, dlp1 = 0.0, dlp2 = 0.0;
    if (fabs(pl) > 1e-18) {
        const double dp = (fa * B.poly + fb * A.poly) * 0.5 * rr / r2;
        dlp0 = dx * (dp / pl);
        dlp1 = dy * (dp / pl);
        dlp2 = dz * (dp / pl);
    }
    const double hp2 = 2.0 * hij * (double)sp;
    const double f0 = (double)g0 + dlp0 * hp2, f1 = (double)g1 + dlp1 * hp2, f2 = (double)g2 + dlp2 * hp2;
    if (!PBC || !self) {
        GradAcc<R>::add(acc.g[0], acc.cg[0], (R)f0);
        GradAcc<R>::add(acc.g[1], acc.cg[1], (R)f1);
        GradAcc<R>::add(acc.g[2], acc.cg[2], (R)f2);
    }
    if (cn_on) GradAcc<R>::add(acc.dc, acc.cdc, (R)(dse_a * hs * (double)sp));
    if constexpr (PBC) {
        // the ket atom's share of a symmetry-unique pair (pbc/gradient.rs:345-384):
        // the opposite force (none for an image of A itself) and its dE/dCN
        if (oth) {
            oth[0] = self ? 0.0 : -f0;
            oth[1] = self ? 0.0 : -f1;
            oth[2] = self ? 0.0 : -f2;
            oth[3] = cn_on ? dse_b * hs * (double)sp : 0.0;
        }
    }
}

// Shell pair (sa, sb) with sb's atom at image (n0, n1, n2), as the bra side of
// atom a (pbc/gradient.rs:285-384 at the Gamma point); the (LA, LB) class is
// chosen by the caller's switch.
template <int LA, int LB, typename R>
__device__ __forceinline__ void grad_pair_image(const DeviceBatch& b, const DeviceTables& t, const double* cn,
                                                const double* vsh, int scc_on, int cn_on, int a, int sa, int sb,
                                                int ld, gptr<const double> P, gptr<const double> W,
                                                const CellRec& cr, int n0, int n1, int n2, GradAcc<R>& acc,
                                                double* oth) {
    double tv[3];
    image_vector(cr, n0, n1, n2, tv);
    grad_pair<LA, LB, R, true>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, acc, tv, b.sh_atom[sb] == a, oth);
}

// Adds the lanes' ket-side shares v (3 force components and dE/dCN) to the LDS
// accumulator accb[4 key + .] of their ket atom key (-1: no share).  The keys
// of a batch are non-decreasing over the lanes, so equal keys are one run: a
// segmented scan over the wavefront (fixed order) leaves each run's sum on its
// last lane, which adds it.  All 64 lanes call this.
__device__ __forceinline__ void ket_shares_add(int key, double v[4], double* accb) {
    const int lane = threadIdx.x;
    for (int off = 1; off < 64; off <<= 1) {
        const int ku = __shfl_up(key, off, 64);
        double vu[4];
        for (int k = 0; k < 4; ++k) vu[k] = __shfl_up(v[k], off, 64);
        if (lane >= off && ku == key)
            for (int k = 0; k < 4; ++k) v[k] += vu[k];
    }
    const int kn = __shfl_down(key, 1, 64);
    if (key >= 0 && (lane == 63 || kn != key))
        for (int k = 0; k < 4; ++k) accb[4 * key + k] += v[k];
}

}  // namespace

// S and H0 (integrals.rs:133-242; hamiltonian.rs:91-134).  One 64-lane
// work-group per listed atom A = atoms[blockIdx.x]; for every shell of A and every angular-momentum class
// the lanes walk the instance's shells of that class on atoms B <= A, so a
// wavefront always evaluates one (LA, LB) specialisation.
template <bool DCLASS, bool PBC>
__global__ void __launch_bounds__(64) k_overlap_h0_t(DeviceBatch b, DeviceTables t, const int* atoms, const double* cn,
                                                     double cutoff) {
    const int a = atoms[blockIdx.x];
    const int inst = b.atom_inst[a];
    // molecular and periodic instances go to separate instantiations
    if (periodic_instance(b, inst) != PBC) return;
    const int ld = b.ld[inst];
    int rg[3] = {0, 0, 0};
    if constexpr (PBC) image_ranges(b.cell[inst], kPbcAoCutoff, rg);
    __shared__ int pq[PBC ? 128 : 1];  // queue of candidates inside the cutoff (periodic)
    gptr<double> S = (gptr<double>)b.S[inst];
    gptr<double> H = (gptr<double>)b.H0[inst];
    const double c2 = cutoff * cutoff;
    const int sa0 = b.atom_sh0[a];
    const int nsha = t.elements[b.atom_z[a]].nshell;
    const int* lo = b.lofs + 4 * inst;
    for (int k = 0; k < nsha; ++k) {
        const int sa = sa0 + k;
        const int la = t.shells[b.sh_type[sa]].l;
        for (int lb = 0; lb < 3; ++lb) {
            // s/p classes and d classes go to separate instantiations (register budget)
            if ((la == 2 || lb == 2) != DCLASS) continue;
            if constexpr (PBC) {
                // the class's shells with index <= sa (a prefix: shells are numbered
                // atom by atom and each class list is in shell order)
                int nvalid = 0;
                for (int base = lo[lb]; base < lo[lb + 1]; base += 64) {
                    const int idx = base + threadIdx.x;
                    nvalid += __popcll(__ballot(idx < lo[lb + 1] && b.sh_by_l[idx] <= sa));
                }
                if (nvalid == 0) continue;
                const CellRec& cr = b.cell[inst];
                if constexpr (!DCLASS) {
                    switch (la * 3 + lb) {
                        case 0: s_h0_shell_pbc<0, 0>(b, t, cn, cr, rg, a, sa, lo[lb], nvalid, ld, S, H, pq); break;
                        case 1: s_h0_shell_pbc<0, 1>(b, t, cn, cr, rg, a, sa, lo[lb], nvalid, ld, S, H, pq); break;
                        case 3: s_h0_shell_pbc<1, 0>(b, t, cn, cr, rg, a, sa, lo[lb], nvalid, ld, S, H, pq); break;
                        default: s_h0_shell_pbc<1, 1>(b, t, cn, cr, rg, a, sa, lo[lb], nvalid, ld, S, H, pq); break;
                    }
                } else {
                    switch (la * 3 + lb) {
                        case 2: s_h0_shell_pbc<0, 2>(b, t, cn, cr, rg, a, sa, lo[lb], nvalid, ld, S, H, pq); break;
                        case 5: s_h0_shell_pbc<1, 2>(b, t, cn, cr, rg, a, sa, lo[lb], nvalid, ld, S, H, pq); break;
                        case 6: s_h0_shell_pbc<2, 0>(b, t, cn, cr, rg, a, sa, lo[lb], nvalid, ld, S, H, pq); break;
                        case 7: s_h0_shell_pbc<2, 1>(b, t, cn, cr, rg, a, sa, lo[lb], nvalid, ld, S, H, pq); break;
                        default: s_h0_shell_pbc<2, 2>(b, t, cn, cr, rg, a, sa, lo[lb], nvalid, ld, S, H, pq); break;
                    }
                }
                continue;
            }
            for (int idx = lo[lb] + threadIdx.x; idx < lo[lb + 1]; idx += blockDim.x) {
                const int sb = b.sh_by_l[idx];
                const int ob = b.sh_atom[sb];
                if (ob > a || (ob == a && sb > sa)) continue;
                if constexpr (!DCLASS) {
                    switch (la * 3 + lb) {
                        case 0: s_h0_pair<0, 0>(b, t, cn, c2, a, sa, sb, ld, S, H); break;
                        case 1: s_h0_pair<0, 1>(b, t, cn, c2, a, sa, sb, ld, S, H); break;
                        case 3: s_h0_pair<1, 0>(b, t, cn, c2, a, sa, sb, ld, S, H); break;
                        default: s_h0_pair<1, 1>(b, t, cn, c2, a, sa, sb, ld, S, H); break;
                    }
                } else {
                    switch (la * 3 + lb) {
                        case 2: s_h0_pair<0, 2>(b, t, cn, c2, a, sa, sb, ld, S, H); break;
                        case 5: s_h0_pair<1, 2>(b, t, cn, c2, a, sa, sb, ld, S, H); break;
                        case 6: s_h0_pair<2, 0>(b, t, cn, c2, a, sa, sb, ld, S, H); break;
                        case 7: s_h0_pair<2, 1>(b, t, cn, c2, a, sa, sb, ld, S, H); break;
                        default: s_h0_pair<2, 2>(b, t, cn, c2, a, sa, sb, ld, S, H); break;
                    }
                }
            }
        }
    }
}

// Electronic gradient, pair part (gradient.rs:333-462) for atom A = atoms[j]:
// every pair with B != A inside the integral cutoff, A taken as the bra.
// Deterministic per-atom sums (one 64-lane wavefront per atom, no atomics).
// PBC: the instantiation for periodic instances (the others return at once);
// images within 30 bohr, the shell potential always on (pbc/gradient.rs:250-384).
// With ket_rows (the symmetry-unique pairs of the reference: the origin cell
// with B > A, every image with positive canonical index with every B,
// pbc/gradient.rs:305-313) each pair is evaluated once, by its bra atom A: A's
// own share goes to g_pair / dedcn_h, the ket atom's share (opposite force,
// its dE/dCN) to row A of the instance's ket-share matrix at ket_rows +
// cell.ket0 (4 doubles per (A, B), summed over A by k_periodic_merge), through
// an LDS accumulator of 4 nat doubles (dynamic shared memory).  Without
// ket_rows every image of every atom is evaluated from both sides.
template <bool DCLASS, bool MIXED, bool PBC>
__global__ void __launch_bounds__(64) k_grad_pairs_t(DeviceBatch b, DeviceTables t, const int* atoms,
                                                     const double* cn, const double* vsh, int scc_on, int cn_on,
                                                     double cutoff, double* g_pair, double* dedcn_h,
                                                     double* ket_rows) {
    const int j = blockIdx.x;
    const int a = atoms[j];
    const int inst = b.atom_inst[a];
    if (periodic_instance(b, inst) != PBC) return;
    const int ld = b.ld[inst];
    int rg[3] = {0, 0, 0};
    __shared__ int pq[PBC ? 128 : 1];  // queue of candidates inside the cutoff (periodic)
    extern __shared__ double accb[];   // ket-share accumulator (periodic, 4 per atom)
    const int a0 = b.atom0[inst], nat = b.atom0[inst + 1] - a0;
    const bool halve = PBC && ket_rows != nullptr;
    if constexpr (PBC) {
        image_ranges(b.cell[inst], kPbcAoCutoff, rg);
        scc_on = 1;
        if (halve) {
            for (int i = threadIdx.x; i < 4 * nat; i += 64) accb[i] = 0.0;
            __syncthreads();
        }
    }
    gptr<const double> P = (gptr<const double>)b.P[inst];
    gptr<const double> W = (gptr<const double>)b.W[inst];
    const double c2 = cutoff * cutoff;
    const int sa0 = b.atom_sh0[a];
    const int nsha = t.elements[b.atom_z[a]].nshell;
    const int* lo = b.lofs + 4 * inst;
    using R = typename std::conditional<MIXED, float, double>::type;
    GradAcc<R> acc;
    for (int k = 0; k < nsha; ++k) {
        const int sa = sa0 + k;
        const int la = t.shells[b.sh_type[sa]].l;
        for (int lb = 0; lb < 3; ++lb) {
            if ((la == 2 || lb == 2) != DCLASS) continue;
            if constexpr (PBC) {
                // Candidates are (shell of class lb, image) pairs, 64 at a time,
                // one per lane; the ones inside the 30-bohr cutoff and outside the
                // overlap screen are compacted into an LDS queue (ballot order, so
                // the order is fixed) and evaluated 64 at a time, one per lane, so
                // the lanes do not idle on images that are out of reach for their
                // own shell.
                const CellRec& cr = b.cell[inst];
                const int d0 = 2 * rg[0] + 1, d1 = 2 * rg[1] + 1, d2 = 2 * rg[2] + 1;
                const int nimg = d0 * d1 * d2, s_first = lo[lb];
                const int ncand = (lo[lb + 1] - s_first) * nimg;
                const int lane = threadIdx.x;
                const double* x = b.atom_x;
                const double c2p = kPbcAoCutoff * kPbcAoCutoff;
                auto decode = [&](int c, int& sb, int& n0, int& n1, int& n2) {
                    sb = b.sh_by_l[s_first + c / nimg];
                    const int im = c % nimg;
                    n0 = im / (d1 * d2) - rg[0];
                    n1 = (im / d2) % d1 - rg[1];
                    n2 = im % d2 - rg[2];
                };
                const int center = (nimg - 1) / 2;  // the origin image
                int key = -1;                          // ket atom of the lane's evaluation
                double ov[4] = {0.0, 0.0, 0.0, 0.0};   // its ket-side share
                double* oth = halve ? ov : nullptr;
                auto evaluate = [&](int c) {
                    int sb, n0, n1, n2;
                    decode(c, sb, n0, n1, n2);
                    key = b.sh_atom[sb] - a0;
                    if constexpr (!DCLASS) {
                        switch (la * 3 + lb) {
                            case 0: grad_pair_image<0, 0, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, cr, n0, n1, n2, acc, oth); break;
                            case 1: grad_pair_image<0, 1, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, cr, n0, n1, n2, acc, oth); break;
                            case 3: grad_pair_image<1, 0, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, cr, n0, n1, n2, acc, oth); break;
                            default: grad_pair_image<1, 1, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, cr, n0, n1, n2, acc, oth); break;
                        }
                    } else {
                        switch (la * 3 + lb) {
                            case 2: grad_pair_image<0, 2, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, cr, n0, n1, n2, acc, oth); break;
                            case 5: grad_pair_image<1, 2, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, cr, n0, n1, n2, acc, oth); break;
                            case 6: grad_pair_image<2, 0, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, cr, n0, n1, n2, acc, oth); break;
                            case 7: grad_pair_image<2, 1, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, cr, n0, n1, n2, acc, oth); break;
                            default: grad_pair_image<2, 2, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, cr, n0, n1, n2, acc, oth); break;
                        }
                    }
                };
                int cursor = 0;  // queued candidates, uniform over the wavefront
                for (int c0 = 0; c0 < ncand; c0 += 64) {
                    const int c = c0 + lane;
                    bool inside = false;
                    if (c < ncand) {
                        int sb, n0, n1, n2;
                        decode(c, sb, n0, n1, n2);
                        const int ob = b.sh_atom[sb];
                        const int im = c % nimg;
                        // symmetry-unique pairs only when the ket shares are kept
                        const bool unique = !halve || im > center || (im == center && ob > a);
                        if (unique && !(ob == a && n0 == 0 && n1 == 0 && n2 == 0)) {
                            double tv[3];
                            image_vector(cr, n0, n1, n2, tv);
                            const double dx = x[3 * a] - x[3 * ob] - tv[0], dy = x[3 * a + 1] - x[3 * ob + 1] - tv[1],
                                         dz = x[3 * a + 2] - x[3 * ob + 2] - tv[2];
                            const double r2 = dx * dx + dy * dy + dz * dz;
                            inside = r2 > 1e-12 && r2 <= c2p && !overlap_screened(t, b.atom_z[a], b.atom_z[ob], r2);
                        }
                    }
                    const unsigned long long m = __ballot(inside);
                    if (inside) pq[cursor + __popcll(m & ((1ull << lane) - 1ull))] = c;
                    cursor += __popcll(m);
                    __syncthreads();
                    if (cursor >= 64) {
                        evaluate(pq[lane]);
                        if (halve) ket_shares_add(key, ov, accb);
                        __syncthreads();
                        if (lane < cursor - 64) pq[lane] = pq[64 + lane];
                        cursor -= 64;
                        __syncthreads();
                    }
                }
                key = -1;
                if (lane < cursor) evaluate(pq[lane]);
                if (halve) ket_shares_add(key, ov, accb);
                __syncthreads();
                continue;
            }
            for (int idx = lo[lb] + threadIdx.x; idx < lo[lb + 1]; idx += blockDim.x) {
                const int sb = b.sh_by_l[idx];
                const int ob = b.sh_atom[sb];
                if (ob == a) continue;
                const double r2 = dist2(b.atom_x, a, ob);
                if (r2 <= kPairEps2 || r2 > c2 || r2 <= 1e-12) continue;
                if constexpr (!DCLASS) {
                    switch (la * 3 + lb) {
                        case 0: grad_pair<0, 0, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, acc); break;
                        case 1: grad_pair<0, 1, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, acc); break;
                        case 3: grad_pair<1, 0, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, acc); break;
                        default: grad_pair<1, 1, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, acc); break;
                    }
                } else {
                    switch (la * 3 + lb) {
                        case 2: grad_pair<0, 2, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, acc); break;
                        case 5: grad_pair<1, 2, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, acc); break;
                        case 6: grad_pair<2, 0, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, acc); break;
                        case 7: grad_pair<2, 1, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, acc); break;
                        default: grad_pair<2, 2, R>(b, t, cn, vsh, scc_on, cn_on, a, sa, sb, ld, P, W, acc); break;
                    }
                }
            }
        }
    }
    double g[3];
    for (int d = 0; d < 3; ++d) g[d] = wave_sum(acc.value(d));
    const double dc = wave_sum(acc.dcn());
    if (threadIdx.x == 0) {
        // the s/p pass initialises, the d pass accumulates
        for (int d = 0; d < 3; ++d) g_pair[3 * j + d] = DCLASS ? g_pair[3 * j + d] + g[d] : g[d];
        dedcn_h[j] = DCLASS ? dedcn_h[j] + dc : dc;
    }
    if (halve) {
        double* row = ket_rows + b.cell[inst].ket0 + (size_t)(a - a0) * 4 * nat;
        for (int i = threadIdx.x; i < 4 * nat; i += 64) row[i] = DCLASS ? row[i] + accb[i] : accb[i];
    }
}

// Periodic atoms' gradient parts computed apart from the per-bucket pair sums:
// the ket-side shares of the symmetry-unique pairs (k_grad_pairs_t with ket
// rows; none without) summed over the bra atoms in their order, and the Ewald
// gradient g_ew; added to the per-atom electronic gradient and dE/dCN.  A
// thread per listed periodic atom.
__global__ void k_periodic_merge(DeviceBatch b, const int* patoms, int npatoms, const double* ket_rows,
                                 const double* g_ew, double* g, double* dedcn) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npatoms) return;
    const int o = patoms[j];
    const int inst = b.atom_inst[o];
    const int a0 = b.atom0[inst], nat = b.atom0[inst + 1] - a0;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if (ket_rows) {
        const double* col = ket_rows + b.cell[inst].ket0 + (size_t)(o - a0) * 4;
        for (int r = 0; r < nat; ++r)
            for (int k = 0; k < 4; ++k) s[k] += col[(size_t)r * 4 * nat + k];
    }
    for (int d = 0; d < 3; ++d) g[3 * o + d] += s[d] + g_ew[3 * o + d];
    dedcn[o] += s[3];
}

__global__ void k_scatter_atoms(const int* atoms, int n, const double* gl, const double* dcl, double* g, double* dc) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int a = atoms[j];
    for (int d = 0; d < 3; ++d) g[3 * a + d] = gl[3 * j + d];
    dc[a] = dcl[j];
}

__global__ void k_gather_state(const int* act, int n, const SccState* st, SccState* out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = st[act[j]];
    __threadfence_system();
}

// Active list of a captured bucket iteration: position j keeps its instance if
// that instance's next iteration is a plain one, else -1 (the iteration's
// kernels skip it).
__global__ void k_gate_active(const int* inst, int count, const SccState* st, int kcap, int ld, int broyden_memory,
                              int* act) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    const int i = inst[j];
    act[j] = plain_iteration(st[i], kcap, ld, broyden_memory) ? i : -1;
}

__global__ void k_copy_words(const unsigned int* src, unsigned int* dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template __global__ void k_overlap_h0_t<false, false>(DeviceBatch, DeviceTables, const int*, const double*, double);
template __global__ void k_overlap_h0_t<true, false>(DeviceBatch, DeviceTables, const int*, const double*, double);
template __global__ void k_overlap_h0_t<false, true>(DeviceBatch, DeviceTables, const int*, const double*, double);
template __global__ void k_overlap_h0_t<true, true>(DeviceBatch, DeviceTables, const int*, const double*, double);
template __global__ void k_grad_pairs_t<false, false, false>(DeviceBatch, DeviceTables, const int*, const double*,
                                                         const double*, int, int, double, double*, double*, double*);
template __global__ void k_grad_pairs_t<false, false, true>(DeviceBatch, DeviceTables, const int*, const double*,
                                                         const double*, int, int, double, double*, double*, double*);
template __global__ void k_grad_pairs_t<false, true, false>(DeviceBatch, DeviceTables, const int*, const double*,
                                                         const double*, int, int, double, double*, double*, double*);
template __global__ void k_grad_pairs_t<false, true, true>(DeviceBatch, DeviceTables, const int*, const double*,
                                                         const double*, int, int, double, double*, double*, double*);
template __global__ void k_grad_pairs_t<true, false, false>(DeviceBatch, DeviceTables, const int*, const double*,
                                                         const double*, int, int, double, double*, double*, double*);
template __global__ void k_grad_pairs_t<true, false, true>(DeviceBatch, DeviceTables, const int*, const double*,
                                                         const double*, int, int, double, double*, double*, double*);
template __global__ void k_grad_pairs_t<true, true, false>(DeviceBatch, DeviceTables, const int*, const double*,
                                                         const double*, int, int, double, double*, double*, double*);
template __global__ void k_grad_pairs_t<true, true, true>(DeviceBatch, DeviceTables, const int*, const double*,
                                                         const double*, int, int, double, double*, double*, double*);

}  // namespace gfn1
