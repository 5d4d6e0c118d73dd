// Cartesian Hessians by central differences of the analytic gradient.
//
// The reference differentiates the gradient analytically (analytic_hessian,
// src/hessian/mod.rs:112-134) and checks that result against exactly this
// construction, H[:, c] = (g(x + h e_c) - g(x - h e_c)) / 2h with tight SCC
// tolerances (tests/hessian.rs:691-760, agreement within 1e-6 Eh/bohr^2).
// Here the 6 nat displaced geometries of every system are independent
// energy + gradient evaluations, so they run as instances of the batch engine,
// a few thousand per device pass.
#include "hessian.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

namespace gfn1 {
namespace {
constexpr double kMatricesPerInstance = 12.0;  // n x n matrices the batch engine keeps per instance
constexpr double kPassBytes = 64.0e9;           // matrices of one pass of displaced geometries
}  // namespace

Status hessian_batch(const DeviceModel& dm, const std::vector<SystemInput>& systems, EngineOptions e,
                     const HessianOptions& opt, std::vector<HessianResult>& out, HessianStats* stats) {
    if (!(opt.step > 0.0) || !std::isfinite(opt.step))
        return Status::error(kErrInvalidInput, "finite-difference step must be positive");
    if (!(opt.scc_energy_tolerance > 0.0) || !(opt.scc_charge_tolerance > 0.0))
        return Status::error(kErrInvalidInput, "SCC tolerances must be positive");
    const auto t0 = std::chrono::steady_clock::now();
    HessianStats st_local;
    e.energy_tolerance = opt.scc_energy_tolerance;
    e.charge_tolerance = opt.scc_charge_tolerance;
    e.compute_gradient = 1;
    e.include_repulsion = opt.include_repulsion;
    e.include_dispersion = opt.include_dispersion;
    e.include_halogen = opt.include_halogen;
    e.include_hamiltonian = opt.include_electronic;
    e.include_scc = opt.include_electronic;
    e.mixed = 0;  // differences of gradients need every digit
    e.keep_matrices = 0;
    const int n = (int)systems.size();
    out.assign(n, HessianResult{});
    if (n == 0) return Status{};

    // the undisplaced geometries: energy, gradient and start charges
    std::vector<std::vector<double>> q0(n);
    std::vector<double> inst_bytes(n, 0.0);  // device matrices of one displaced copy
    {
        Batch base;
        Status st = base.create(dm, systems, e);
        if (!st.ok()) return st;
        st = base.run();
        if (!st.ok()) return st;
        std::vector<int> status;
        std::vector<double> energy, grad, q, dip;
        st = base.fetch_energy_gradient(status, energy, grad, &q);
        if (!st.ok()) return st;
        st = base.fetch_dipoles(dip);
        if (!st.ok()) return st;
        ++st_local.passes;
        st_local.evaluations += n;
        for (int i = 0; i < n; ++i) {
            HessianResult& r = out[i];
            const double ld = base.padded_dimension(i);
            inst_bytes[i] = kMatricesPerInstance * ld * ld * sizeof(double);
            if (systems[i].nat == 0) {
                r.status = kErrInvalidInput;
                r.message = "cannot build the Hessian of an empty system";
                continue;
            }
            r.status = status[i];
            if (status[i] != kOk) {
                r.message = base.message(i);
                continue;
            }
            const int a0 = base.atom_offset(i), nat = systems[i].nat;
            r.energy = energy[i];
            for (int d = 0; d < 3; ++d) r.dipole[d] = dip[3 * (size_t)i + d];
            r.gradient.assign(grad.begin() + 3 * (size_t)a0, grad.begin() + 3 * (size_t)(a0 + nat));
            q0[i].assign(q.begin() + base.shell_offset(i), q.begin() + base.shell_offset(i) + base.shell_count(i));
        }
    }

    // displaced geometries, in passes of at most max_batch instances
    struct Disp {
        int sys, col, sign;
    };
    std::vector<Disp> all;
    for (int i = 0; i < n; ++i)
        if (out[i].status == kOk) {
            out[i].hessian.assign((size_t)9 * systems[i].nat * systems[i].nat, 0.0);
            out[i].h_repulsion.assign(out[i].hessian.size(), 0.0);
            out[i].h_dispersion.assign(out[i].hessian.size(), 0.0);
            out[i].h_halogen.assign(out[i].hessian.size(), 0.0);
            out[i].dipole_derivatives.assign((size_t)9 * systems[i].nat, 0.0);
            for (int c = 0; c < 3 * systems[i].nat; ++c) {
                all.push_back({i, c, +1});
                all.push_back({i, c, -1});
            }
        }
    // a pass holds at most max_batch displaced geometries and about
    // kPassBytes of matrices, and always whole +/- pairs
    const size_t chunk = (size_t)std::max(2, opt.max_batch) & ~(size_t)1;
    std::vector<char> failed(n, 0);
    for (size_t s0 = 0, m = 0; s0 < all.size(); s0 += m) {
        double bytes = 0.0;
        m = 0;
        while (s0 + m < all.size() && m < chunk) {
            const double pair = 2.0 * inst_bytes[all[s0 + m].sys];
            if (m > 0 && bytes + pair > kPassBytes) break;
            bytes += pair;
            m += 2;
        }
        std::vector<std::vector<double>> pos(m);
        std::vector<SystemInput> in(m);
        for (size_t k = 0; k < m; ++k) {
            const Disp& d = all[s0 + k];
            const SystemInput& sys = systems[d.sys];
            pos[k].assign(sys.x, sys.x + 3 * (size_t)sys.nat);
            pos[k][d.col] += d.sign * opt.step;
            in[k] = sys;
            in[k].x = pos[k].data();
            if (opt.warm_start) {
                in[k].q0 = q0[d.sys].data();
                in[k].q0_len = (int)q0[d.sys].size();
            }
        }
        Batch b;
        Status st = b.create(dm, in, e);
        if (!st.ok()) return st;
        st = b.run();
        if (!st.ok()) return st;
        std::vector<int> status;
        std::vector<double> energy, grad, dip, terms;
        st = b.fetch_energy_gradient(status, energy, grad, nullptr, &terms);
        if (!st.ok()) return st;
        st = b.fetch_dipoles(dip);
        if (!st.ok()) return st;
        ++st_local.passes;
        st_local.evaluations += (long long)m;
        // entries k (+h) and k + 1 (-h) of a pass are the two sides of one column
        for (size_t k = 0; k + 1 < m; k += 2) {
            const Disp& d = all[s0 + k];
            HessianResult& r = out[d.sys];
            const int bad = status[k] != kOk ? (int)k : (status[k + 1] != kOk ? (int)k + 1 : -1);
            if (bad >= 0) {
                if (!failed[d.sys]) {
                    failed[d.sys] = 1;
                    r.status = status[bad];
                    r.message = b.message(bad);
                }
                continue;
            }
            const int ndof = 3 * systems[d.sys].nat;
            const double* gp = grad.data() + 3 * (size_t)b.atom_offset((int)k);
            const double* gm = grad.data() + 3 * (size_t)b.atom_offset((int)k + 1);
            for (int row = 0; row < ndof; ++row)
                r.hessian[(size_t)row * ndof + d.col] = (gp[row] - gm[row]) / (2.0 * opt.step);
            // the classical terms' parts from their own gradients (the reference's
            // repulsion, dispersion and halogen sub-results, src/hessian/mod.rs:55-65)
            std::vector<double>* parts[3] = {&r.h_repulsion, &r.h_dispersion, &r.h_halogen};
            const size_t nb = (size_t)b.natoms();
            for (int t = 0; t < 3; ++t) {
                const double* tp = terms.data() + 3 * (t * nb + (size_t)b.atom_offset((int)k));
                const double* tm = terms.data() + 3 * (t * nb + (size_t)b.atom_offset((int)k + 1));
                for (int row = 0; row < ndof; ++row)
                    (*parts[t])[(size_t)row * ndof + d.col] = (tp[row] - tm[row]) / (2.0 * opt.step);
            }
            // dipole derivatives by the same differences (the analytic route is
            // src/properties.rs:97-145)
            for (int a = 0; a < 3; ++a)
                r.dipole_derivatives[3 * (size_t)d.col + a] =
                    (dip[3 * k + a] - dip[3 * (k + 1) + a]) / (2.0 * opt.step);
        }
    }
    for (int i = 0; i < n; ++i) {
        HessianResult& r = out[i];
        if (r.status != kOk) {
            r.hessian.clear();
            r.dipole_derivatives.clear();
            r.h_repulsion.clear();
            r.h_dispersion.clear();
            r.h_halogen.clear();
            continue;
        }
        const int ndof = 3 * systems[i].nat;
        auto symmetrise = [&](std::vector<double>& h) {
            double asym = 0.0;
            for (int a = 0; a < ndof; ++a)
                for (int c = 0; c < a; ++c) {
                    double& x = h[(size_t)a * ndof + c];
                    double& y = h[(size_t)c * ndof + a];
                    asym = std::max(asym, std::fabs(x - y));
                    const double avg = 0.5 * (x + y);
                    x = y = avg;
                }
            return asym;
        };
        r.max_asymmetry = symmetrise(r.hessian);
        symmetrise(r.h_repulsion);
        symmetrise(r.h_dispersion);
        symmetrise(r.h_halogen);
    }
    st_local.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st_local;
    return Status{};
}

}  // namespace gfn1
