// Periodic Gamma-point path: the Ewald-summed shell Coulomb matrix and its
// nuclear gradient (reference: src/pbc/ewald.rs, src/pbc/gradient.rs:548-779),
// and the host-side cell records (src/lattice.rs).
//
// The Klopman-Ohno interaction gamma(R) = 1/sqrt(R^2 + eta^-2) is split as
//   gamma = [1/R - eta^-2 / (2 R^3)] + [gamma - 1/R + eta^-2 / (2 R^3)]:
// the 1/R part by a standard Ewald sum over atom centres, the R^-3 part by the
// generalised Ewald sum (real space with erfc, reciprocal space with the
// exponential integral E1, the k = 0 logarithm and the self term), and the
// remainder, which decays as R^-5, directly in real space within 10 bohr.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "periodic.hpp"

namespace gfn1 {

namespace {

constexpr double kPiP = 3.14159265358979323846;
constexpr double kSqrtPi = 1.7724538509055160273;
constexpr double kR3Coeff = -0.5;  // binomial coefficient of the eta^-2 R^-3 term (pbc/ewald.rs:38)

__device__ __forceinline__ double sum3(double x, double y, double z) { return x * x + y * y + z * z; }

// Real-space part of the eta^-2 R^-3 Ewald term and its radial derivative
// (pbc/ewald.rs:121-140), without the eta^-2 factor.
__device__ __forceinline__ void r3_real(double r, double alpha, double& v, double& dv) {
    const double ar = alpha * r, ex = exp(-ar * ar);
    const double q = erfc(ar) + (2.0 * ar / kSqrtPi) * ex;
    const double dq = -(4.0 * alpha * alpha * alpha * r * r / kSqrtPi) * ex;
    const double r3 = r * r * r;
    v = q / r3;
    dv = dq / r3 - 3.0 * q / (r3 * r);
}

// Short-range remainder gamma - 1/R + eta^-2 / (2 R^3) and its radial
// derivative (pbc/ewald.rs:108-119).
__device__ __forceinline__ void ko_short(double r, double ie2, double& v, double& dv) {
    const double den2 = r * r + ie2, den = sqrt(den2);
    const double r2 = r * r, r3 = r2 * r;
    v = 1.0 / den - 1.0 / r + 0.5 * ie2 / r3;
    dv = -r / (den2 * den) + 1.0 / r2 - 1.5 * ie2 / (r2 * r2);
}

__device__ __forceinline__ void image_ranges(const CellRec& c, double cut, int r[3]) {
    for (int k = 0; k < 3; ++k) r[k] = (int)ceil(cut / c.h[k]) + 1;
}

}  // namespace

// Decodes p = la (la + 1) / 2 + lb, lb <= la (the lower triangle row by row).
__device__ __forceinline__ void triangle_pair(int p, int& la, int& lb) {
    la = (int)((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
    while (la * (la + 1) / 2 > p) --la;
    while ((la + 1) * (la + 2) / 2 <= p) ++la;
    lb = p - la * (la + 1) / 2;
}

// Gamma_ij of every periodic instance (pbc/ewald.rs:200-320).  A thread per
// atom pair (a, b <= a) of an instance, 64 pairs per work-group, the instance
// in blockIdx.y.  The thread walks the lattice images of its pair once: the
// parts that depend on the distance only (erfc / R, the R^-3 real-space factor,
// the reciprocal sums) are summed per atom pair, the short-range remainder per
// shell pair (it depends on the pair's eta; the at most 3 x 3 sums sit in an
// LDS column of the thread), and the thread writes Gamma_ij and Gamma_ji for the shells
// of a and b (the matrix is exactly symmetric).  Images are tested on R^2
// first; the inclusion tests themselves are the reference's, on R.
__global__ void __launch_bounds__(64) k_ewald_gamma(DeviceBatch b, DeviceTables t, const int* pinst) {
    const int inst = pinst[blockIdx.y];
    const int a0 = b.atom0[inst], nat = b.atom0[inst + 1] - a0;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nat * (nat + 1) / 2) return;
    int la, lb;
    triangle_pair(p, la, lb);
    const CellRec& c = b.cell[inst];
    const double alpha = c.alpha;
    const double rcut = kEwaldTau / alpha;
    const double realcut = fmax(rcut, kEwaldShortCutoff);
    // R^2 above this is certainly outside realcut after the square root
    const double pre2 = realcut * realcut * (1.0 + 1e-12);
    int rg[3];
    image_ranges(c, realcut, rg);
    const int s0 = b.sh0[inst], nsh = b.nsh[inst];
    const double* x = b.atom_x;
    const double* gt = b.gtab + c.g0 * kGStride;
    double* gam = const_cast<double*>(b.gam) + c.gam0;
    const int A = a0 + la, B = a0 + lb;
    const int fa = b.atom_sh0[A], na = t.elements[b.atom_z[A]].nshell;
    const int fb = b.atom_sh0[B], nb = t.elements[b.atom_z[B]].nshell;
    const bool same = A == B;
    constexpr int M = kMaxShellsPerAtom;
    // eta^-2 and the remainder sums of the shell pairs, one column per lane
    __shared__ double ie2[M * M][64], sr[M * M][64];
    const int lane = threadIdx.x;
    const int npq = na * nb;
    for (int k = 0; k < npq; ++k) {
        const double hi = t.shells[b.sh_type[fa + k / nb]].hardness, hj = t.shells[b.sh_type[fb + k % nb]].hardness;
        const double eta = 2.0 / (1.0 / hi + 1.0 / hj);
        ie2[k][lane] = 1.0 / (eta * eta);
        sr[k][lane] = 0.0;
    }
    const double rab[3] = {x[3 * A] - x[3 * B], x[3 * A + 1] - x[3 * B + 1], x[3 * A + 2] - x[3 * B + 2]};
    double real1 = 0.0, real3 = 0.0;
    for (int n0 = -rg[0]; n0 <= rg[0]; ++n0)
        for (int n1 = -rg[1]; n1 <= rg[1]; ++n1)
            for (int n2 = -rg[2]; n2 <= rg[2]; ++n2) {
                double d[3];
                for (int k = 0; k < 3; ++k) d[k] = rab[k] - (n0 * c.a[0][k] + n1 * c.a[1][k] + n2 * c.a[2][k]);
                const double r2 = sum3(d[0], d[1], d[2]);
                if (r2 > pre2) continue;
                const double r = sqrt(r2);
                // r <= 1e-12 is the atom itself in the origin cell
                if (r <= 1e-12 || r > realcut) continue;
                if (r <= rcut) {
                    const double ar = alpha * r, ec = erfc(ar), r3 = r * r * r;
                    real1 += ec / r;
                    real3 += (ec + (2.0 * ar / kSqrtPi) * exp(-ar * ar)) / r3;
                }
                if (r <= kEwaldShortCutoff)
                    for (int k = 0; k < npq; ++k) {
                        double v, dv;
                        ko_short(r, ie2[k][lane], v, dv);
                        sr[k][lane] += v;
                    }
            }
    double rec1 = 0.0, rec3 = 0.0;
    for (int g = 0; g < c.ng; ++g) {
        const double* e = gt + (size_t)g * kGStride;
        const double cs = cos(e[0] * rab[0] + e[1] * rab[1] + e[2] * rab[2]);
        rec1 += e[3] * cs;
        rec3 += e[4] * cs;
    }
    const double phi = real1 + rec1 + c.ew_bg + (same ? c.ew_self : 0.0);
    for (int k = 0; k < npq; ++k) {
        const int p1 = k / nb, q1 = k % nb;
        const int i = fa + p1 - s0, j = fb + q1 - s0;
        if (j > i) continue;  // the same atom: lower triangle of its block
        const double hi = t.shells[b.sh_type[fa + p1]].hardness, hj = t.shells[b.sh_type[fb + q1]].hardness;
        const double eta = 2.0 / (1.0 / hi + 1.0 / hj);
        const double e2 = ie2[k][lane];
        double v = phi + (same ? eta : 0.0);
        v += kR3Coeff * e2 * real3 + e2 * rec3 + c.r3_k0 * e2;
        if (same) v += c.r3_self * e2;
        v += sr[k][lane];
        gam[(size_t)i * nsh + j] = v;
        gam[(size_t)j * nsh + i] = v;
    }
}

// Charge structure factors of every periodic instance at the converged shell
// charges (the reciprocal sums of pbc/gradient.rs:679-779 factorised): a thread
// per (instance, G), the shells of the instance summed in their order.
__global__ void __launch_bounds__(64) k_ewald_sfac(DeviceBatch b, DeviceTables t, const int* pinst, const double* q) {
    const int inst = pinst[blockIdx.y];
    const CellRec& c = b.cell[inst];
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= c.ng) return;
    const double* e = b.gtab + (c.g0 + g) * kGStride;
    const int a0 = b.atom0[inst], a1 = b.atom0[inst + 1];
    const double* x = b.atom_x;
    double cs[3] = {0.0, 0.0, 0.0}, sn[3] = {0.0, 0.0, 0.0};
    for (int o = a0; o < a1; ++o) {
        double so, co;
        sincos(e[0] * x[3 * o] + e[1] * x[3 * o + 1] + e[2] * x[3 * o + 2], &so, &co);
        const int f = b.atom_sh0[o], ns = t.elements[b.atom_z[o]].nshell;
        for (int k = 0; k < ns; ++k) {
            const double qj = q[f + k], ih = 1.0 / t.shells[b.sh_type[f + k]].hardness;
            const double w[3] = {qj, qj * ih, qj * ih * ih};
            for (int m = 0; m < 3; ++m) {
                cs[m] += w[m] * co;
                sn[m] += w[m] * so;
            }
        }
    }
    double* out = b.gsf + (c.g0 + g) * kGsfStride;
    for (int m = 0; m < 3; ++m) {
        out[m] = cs[m];
        out[3 + m] = sn[m];
    }
}

// Gradient of 1/2 q^T Gamma q at the converged shell charges for every atom of
// a periodic instance (pbc/gradient.rs:548-603, :679-779), written to g_ew.  One
// 64-lane work-group per listed atom a.  Real space: a work item is one (plane
// n0 of images, atom o) pair, o fastest over the lanes; it walks the images of
// its plane, the erfc / R force with the atomic charges, the R^-3 and remainder
// forces with the shell pairs of a and o (their charge products and eta^-2 in
// registers).  Reciprocal space: a lane per G over the structure factors of
// k_ewald_sfac, with 1/eta_ij^2 = (1/h_i + 1/h_j)^2 / 4 expanded so that the
// shell sum over j factorises.  Fixed lane assignment and a fixed reduction
// order keep the result deterministic.
__global__ void __launch_bounds__(64) k_ewald_grad(DeviceBatch b, DeviceTables t, const int* patoms,
                                                   const double* q, double* g_ew) {
    const int a = patoms[blockIdx.x];
    const int lane = threadIdx.x;
    const int inst = b.atom_inst[a];
    const CellRec& c = b.cell[inst];
    const double alpha = c.alpha;
    const double rcut = kEwaldTau / alpha, realcut = fmax(rcut, kEwaldShortCutoff);
    const double pre2 = realcut * realcut * (1.0 + 1e-12);
    const double* x = b.atom_x;
    const int a0 = b.atom0[inst], a1 = b.atom0[inst + 1];
    const int fa = b.atom_sh0[a], nsa = t.elements[b.atom_z[a]].nshell;
    constexpr int M = kMaxShellsPerAtom;
    auto qatom = [&](int o) {
        const int f = b.atom_sh0[o], ns = t.elements[b.atom_z[o]].nshell;
        double s = 0.0;
        for (int k = 0; k < ns; ++k) s += q[f + k];
        return s;
    };
    const double qa = qatom(a);
    int rg[3];
    image_ranges(c, realcut, rg);
    const int nat = a1 - a0, d0 = 2 * rg[0] + 1;
    double acc[3] = {0.0, 0.0, 0.0};
    const double tas = 2.0 * alpha / kSqrtPi;
    // charge products and eta^-2 of the shell pairs (a, o), one column per lane
    __shared__ double qq[M * M][64], ie2[M * M][64];
    for (int w = lane; w < nat * d0; w += 64) {
        const int n0 = w / nat - rg[0], o = a0 + w % nat;
        const int fo = b.atom_sh0[o], nso = t.elements[b.atom_z[o]].nshell;
        const double qq1 = qa * qatom(o);
        const int npq = nsa * nso;
        double s3 = 0.0;
        for (int k = 0; k < npq; ++k) {
            const int k2 = k / nso, k3 = k % nso;
            const double hi = t.shells[b.sh_type[fa + k2]].hardness, hj = t.shells[b.sh_type[fo + k3]].hardness;
            const double eta = 2.0 / (1.0 / hi + 1.0 / hj);
            ie2[k][lane] = 1.0 / (eta * eta);
            qq[k][lane] = q[fa + k2] * q[fo + k3];
            s3 += q[fa + k2] * q[fo + k3] / (eta * eta);
        }
        for (int n1 = -rg[1]; n1 <= rg[1]; ++n1)
            for (int n2 = -rg[2]; n2 <= rg[2]; ++n2) {
                double d[3];
                for (int k = 0; k < 3; ++k)
                    d[k] = x[3 * a + k] - x[3 * o + k] - (n0 * c.a[0][k] + n1 * c.a[1][k] + n2 * c.a[2][k]);
                const double r2 = sum3(d[0], d[1], d[2]);
                if (r2 > pre2) continue;
                const double r = sqrt(r2);
                // r <= 1e-12 is the atom itself in the origin cell
                if (r <= 1e-12 || r > realcut) continue;
                double rad = 0.0;
                if (r <= rcut) {
                    const double ar = alpha * r, ex = exp(-ar * ar), ec = erfc(ar);
                    // 1/R Ewald force with the atomic charges
                    rad += qq1 * (-ec / (r * r) - tas * ex / r);
                    // R^-3 real-space force: d/dr [q(r) / r^3] times sum_ij q_i q_j eta_ij^-2
                    const double qq3 = ec + (2.0 * ar / kSqrtPi) * ex;
                    const double dq3 = -(4.0 * alpha * alpha * alpha * r * r / kSqrtPi) * ex;
                    const double dd3 = dq3 / (r * r * r) - 3.0 * qq3 / (r * r * r * r);
                    rad += kR3Coeff * dd3 * s3;
                }
                if (r <= kEwaldShortCutoff)
                    for (int k = 0; k < npq; ++k) {
                        double v, dv;
                        ko_short(r, ie2[k][lane], v, dv);
                        rad += qq[k][lane] * dv;
                    }
                for (int k = 0; k < 3; ++k) acc[k] += d[k] * (rad / r);
            }
    }
    // reciprocal 1/R force (atomic charges) and reciprocal R^-3 force (shell
    // charges, coefficient -w3 per G: pbc/gradient.rs:686-706)
    const double* gt = b.gtab + c.g0 * kGStride;
    const double* sf = b.gsf + c.g0 * kGsfStride;
    for (int g = lane; g < c.ng; g += 64) {
        const double* e = gt + (size_t)g * kGStride;
        const double* f = sf + (size_t)g * kGsfStride;
        double sa, ca;
        sincos(e[0] * x[3 * a] + e[1] * x[3 * a + 1] + e[2] * x[3 * a + 2], &sa, &ca);
        // sum_j w_j sin(phi_a - phi_j) for the three weights
        const double x1 = sa * f[0] - ca * f[3], x2 = sa * f[1] - ca * f[4], x3 = sa * f[2] - ca * f[5];
        double s = -e[3] * qa * x1;
        for (int k2 = 0; k2 < nsa; ++k2) {
            const int si = fa + k2;
            const double ih = 1.0 / t.shells[b.sh_type[si]].hardness;
            s += -e[4] * q[si] * 0.25 * (ih * ih * x1 + 2.0 * ih * x2 + x3);
        }
        for (int k = 0; k < 3; ++k) acc[k] += e[k] * s;
    }
    for (int k = 0; k < 3; ++k)
        for (int off = 32; off > 0; off >>= 1) acc[k] += __shfl_down(acc[k], off, 64);
    if (lane == 0)
        for (int k = 0; k < 3; ++k) g_ew[3 * a + k] = acc[k];
}

// ---------------------------------------------------------------- host side

// Exponential integral E1(x), x > 0: its power series
//   E1 = -gamma - ln x + sum_{k>=1} (-1)^(k+1) x^k / (k k!)
// up to x = 1, above that the continued fraction
//   E1 = e^-x / (x + 1 - 1^2 / (x + 3 - 2^2 / (x + 5 - ...)))
// evaluated by the modified Lentz method (Abramowitz & Stegun 5.1.11, 5.1.22).
double exponential_integral_e1(double x) {
    if (x <= 1.0) {
        double term = 1.0, sum = 0.0;
        for (int k = 1; k <= 80; ++k) {
            term *= -x / k;
            const double add = term / k;
            sum += add;
            if (std::fabs(add) < 1e-17 * std::fabs(sum)) break;
        }
        return -0.57721566490153286061 - std::log(x) - sum;
    }
    const double tiny = 1e-300;
    double bq = x + 1.0, cq = 1.0 / tiny, dq = 1.0 / bq, h = dq;
    for (int i = 1; i <= 400; ++i) {
        const double an = -(double)i * i;
        bq += 2.0;
        dq = an * dq + bq;
        if (std::fabs(dq) < tiny) dq = tiny;
        cq = bq + an / cq;
        if (std::fabs(cq) < tiny) cq = tiny;
        dq = 1.0 / dq;
        const double del = cq * dq;
        h *= del;
        if (std::fabs(del - 1.0) < 1e-16) break;
    }
    return h * std::exp(-x);
}

namespace {
void cross(const double* u, const double* v, double* w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}
double norm(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
}  // namespace

bool make_cell_record(const double* cell, CellRec& c, std::vector<double>& gtab) {
    std::memset(&c, 0, sizeof(c));
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            c.a[i][k] = cell[3 * i + k];
            if (!std::isfinite(c.a[i][k])) return false;
        }
    double bc[3], ca[3], ab[3];
    cross(c.a[1], c.a[2], bc);
    cross(c.a[2], c.a[0], ca);
    cross(c.a[0], c.a[1], ab);
    const double det = c.a[0][0] * bc[0] + c.a[0][1] * bc[1] + c.a[0][2] * bc[2];
    // Lattice::new (lattice.rs:73-87): |det| below 1e-14 is a singular cell
    if (!(std::fabs(det) >= 1e-14)) return false;
    c.vol = std::fabs(det);
    c.h[0] = c.vol / std::max(norm(bc), 1e-12);
    c.h[1] = c.vol / std::max(norm(ca), 1e-12);
    c.h[2] = c.vol / std::max(norm(ab), 1e-12);
    // Ewald constants (pbc/ewald.rs:48-60, :200-216, :253-281)
    const double alpha = kSqrtPi / std::cbrt(std::max(c.vol, 1e-6));
    c.alpha = alpha;
    c.ew_self = -2.0 * alpha / kSqrtPi;
    c.ew_bg = -kPiP / (alpha * alpha * c.vol);
    const double kk = alpha / kSqrtPi;
    const double digamma_3_2 = 0.03648997397857652;  // psi(3/2) = 2 - gamma - 2 ln 2
    c.r3_k0 = kR3Coeff * 4.0 * kPiP * (std::log(kk) + 0.5 * (std::log(kPiP) - digamma_3_2)) / c.vol;
    c.r3_self = kR3Coeff * (-(4.0 * kPiP / 3.0) * kk * kk * kk);
    // reciprocal vectors 0 < |G| <= 2 alpha tau (lattice.rs:312-352): index
    // ranges from the plane heights of the reciprocal basis
    double rec[3][3];
    for (int k = 0; k < 3; ++k) {
        rec[0][k] = 2.0 * kPiP * bc[k] / det;
        rec[1][k] = 2.0 * kPiP * ca[k] / det;
        rec[2][k] = 2.0 * kPiP * ab[k] / det;
    }
    double rbc[3], rca[3], rab[3];
    cross(rec[1], rec[2], rbc);
    cross(rec[2], rec[0], rca);
    cross(rec[0], rec[1], rab);
    const double rvol = std::fabs(rec[0][0] * rbc[0] + rec[0][1] * rbc[1] + rec[0][2] * rbc[2]);
    const double rh[3] = {rvol / std::max(norm(rbc), 1e-12), rvol / std::max(norm(rca), 1e-12),
                          rvol / std::max(norm(rab), 1e-12)};
    const double gcut = 2.0 * alpha * kEwaldTau, gc2 = gcut * gcut;
    int rg[3];
    for (int k = 0; k < 3; ++k) rg[k] = (int)std::ceil(gcut / std::max(rh[k], 1e-12)) + 1;
    const double inv4a2 = 1.0 / (4.0 * alpha * alpha);
    c.g0 = (int64_t)(gtab.size() / kGStride);
    int ng = 0;
    for (int h = -rg[0]; h <= rg[0]; ++h)
        for (int k = -rg[1]; k <= rg[1]; ++k)
            for (int l = -rg[2]; l <= rg[2]; ++l) {
                if (h == 0 && k == 0 && l == 0) continue;
                double g[3];
                for (int d = 0; d < 3; ++d) g[d] = h * rec[0][d] + k * rec[1][d] + l * rec[2][d];
                const double g2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
                if (g2 > gc2 + 64.0 * 2.220446049250313e-16 * std::max(gc2, 1.0)) continue;
                gtab.push_back(g[0]);
                gtab.push_back(g[1]);
                gtab.push_back(g[2]);
                gtab.push_back(4.0 * kPiP / c.vol * std::exp(-g2 * inv4a2) / g2);
                gtab.push_back(kR3Coeff * 2.0 * kPiP / c.vol * exponential_integral_e1(g2 * inv4a2));
                ++ng;
            }
    c.ng = ng;
    c.on = 1;
    return true;
}

hipError_t launch_ewald_gamma(hipStream_t s, const DeviceBatch& b, const DeviceTables& t, const int* pinst,
                              int npinst, int max_nat) {
    if (npinst == 0 || max_nat == 0) return hipSuccess;
    const int max_pairs = max_nat * (max_nat + 1) / 2;
    hipLaunchKernelGGL(k_ewald_gamma, dim3((max_pairs + 63) / 64, npinst), dim3(64), 0, s, b, t, pinst);
    return hipGetLastError();
}

hipError_t launch_ewald_grad(hipStream_t s, const DeviceBatch& b, const DeviceTables& t, const int* pinst,
                             int npinst, int max_ng, const int* patoms, int npatoms, const double* q, double* g_ew) {
    if (npatoms == 0 || npinst == 0) return hipSuccess;
    if (max_ng > 0) {
        hipLaunchKernelGGL(k_ewald_sfac, dim3((max_ng + 63) / 64, npinst), dim3(64), 0, s, b, t, pinst, q);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_ewald_grad, dim3(npatoms), dim3(64), 0, s, b, t, patoms, q, g_ew);
    return hipGetLastError();
}

}  // namespace gfn1
