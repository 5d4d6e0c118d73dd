// Batched L-BFGS geometry optimisation (host side).  The energy and gradient
// evaluations of all systems run together on the device through the batch
// engine; the L-BFGS bookkeeping of each system (a few vectors of 3 * nat) is
// host work between the batch runs.
#include "optimizer.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

namespace gfn1 {
namespace {

constexpr double kArmijo = 1.0e-4;  // sufficient-decrease constant (src/optimizer.rs:234)
constexpr int kLineSearchTrials = 24;  // step halvings (src/optimizer.rs:236)

// Sums in index order, as the reference's iterator sums (src/optimizer.rs:325-327).
double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t k = 0; k < a.size(); ++k) s += a[k] * b[k];
    return s;
}

bool all_finite(const std::vector<double>& v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

enum Phase { kStart = 0, kTrial = 1, kDone = 2, kFailed = 3 };

// L-BFGS state of one system (src/optimizer.rs:91-185).
struct System {
    int nat = 0, a0 = 0, nsh = 0;
    Phase phase = kStart;
    std::vector<double> x, g, d, trial, q;  // q: shell charges at x
    double e = 0.0, maxg = 0.0, dd = 0.0, alpha = 0.0;
    int tries = 0;
    int last_err = kOk;
    std::string last_msg;
    std::vector<std::vector<double>> s_hist, y_hist;
    std::vector<double> rho_hist;
};

}  // namespace

// Two-loop recursion (src/optimizer.rs:259-287).
std::vector<double> lbfgs_direction(const std::vector<double>& g, const std::vector<std::vector<double>>& s,
                                    const std::vector<std::vector<double>>& y, const std::vector<double>& rho) {
    std::vector<double> q(g);
    if (s.empty()) {
        for (double& v : q) v = -v;
        return q;
    }
    const size_t m = s.size();
    std::vector<double> alpha(m, 0.0);
    for (size_t i = m; i-- > 0;) {
        alpha[i] = rho[i] * dot(s[i], q);
        for (size_t k = 0; k < q.size(); ++k) q[k] += -alpha[i] * y[i][k];
    }
    const double yy = dot(y[m - 1], y[m - 1]);
    const double ys = dot(y[m - 1], s[m - 1]);
    const double gamma = yy > 1.0e-18 ? ys / yy : 1.0;
    for (double& v : q) v *= gamma;
    for (size_t i = 0; i < m; ++i) {
        const double beta = rho[i] * dot(y[i], q);
        for (size_t k = 0; k < q.size(); ++k) q[k] += (alpha[i] - beta) * s[i][k];
    }
    for (double& v : q) v = -v;
    return q;
}

// Largest per-atom displacement capped at max_atom_step (src/optimizer.rs:290-305).
// fmax ignores a NaN operand, as f64::max does.
void limit_atom_step(std::vector<double>& d, double max_atom_step) {
    if (max_atom_step <= 0.0) return;
    double mx = 0.0;
    for (size_t k = 0; k + 2 < d.size(); k += 3)
        mx = std::fmax(mx, std::sqrt(d[k] * d[k] + d[k + 1] * d[k + 1] + d[k + 2] * d[k + 2]));
    if (mx > max_atom_step) {
        const double scale = max_atom_step / mx;
        for (double& v : d) v *= scale;
    }
}

// Largest Cartesian gradient component (src/gradient.rs:227-230).
double max_gradient(const std::vector<double>& g) {
    double m = 0.0;
    for (size_t k = 0; k + 2 < g.size(); k += 3)
        m = std::fmax(m, std::fmax(std::fmax(std::fabs(g[k]), std::fabs(g[k + 1])), std::fabs(g[k + 2])));
    return m;
}

Status optimize_batch(Batch& batch, const OptimizerOptions& opt, std::vector<OptimizerResult>& out,
                      OptimizerStats* stats) {
    if (opt.max_iterations < 0) return Status::error(kErrInvalidInput, "max_iterations must not be negative");
    const auto t0 = std::chrono::steady_clock::now();
    const int n = batch.size();
    const size_t history_len = (size_t)std::max(opt.history, 1);
    const double alpha0 = std::max(opt.initial_step, 1.0e-8);
    out.assign(n, OptimizerResult{});
    std::vector<System> sys(n);
    std::vector<double> xall(3 * (size_t)batch.natoms());
    if (!xall.empty()) {
        Status st = batch.positions(xall.data());
        if (!st.ok()) return st;
    }
    for (int i = 0; i < n; ++i) {
        System& s = sys[i];
        s.a0 = batch.atom_offset(i);
        s.nat = batch.atom_offset(i + 1) - s.a0;
        s.nsh = batch.shell_count(i);
        s.x.assign(xall.begin() + 3 * (size_t)s.a0, xall.begin() + 3 * (size_t)(s.a0 + s.nat));
        if (s.nat == 0) {
            // src/optimizer.rs:86-90
            s.phase = kFailed;
            out[i].status = kErrInvalidInput;
            out[i].message = "cannot optimize an empty system";
        }
    }
    auto record = [&](int i, int iteration, double step_norm) {
        if (!opt.keep_trajectory) return;
        OptimizerStep step;
        step.iteration = iteration;
        step.energy = sys[i].e;
        step.max_gradient = sys[i].maxg;
        step.step_norm = step_norm;
        step.positions = sys[i].x;
        out[i].trajectory.push_back(std::move(step));
    };
    // next trial of the current line search; a non-finite trial counts as a
    // failed evaluation without a batch run
    auto set_trial = [&](int i) {
        System& s = sys[i];
        while (s.tries < kLineSearchTrials) {
            s.trial.resize(s.x.size());
            for (size_t k = 0; k < s.x.size(); ++k) s.trial[k] = s.x[k] + s.alpha * s.d[k];
            if (all_finite(s.trial)) {
                s.phase = kTrial;
                return;
            }
            s.last_err = kErrInvalidInput;
            s.last_msg = "non-finite coordinate";
            s.alpha *= 0.5;
            ++s.tries;
        }
        s.phase = kFailed;
        out[i].status = s.last_err;
        out[i].message = s.last_msg;
    };
    // start of the next iteration, or the end (src/optimizer.rs:117-130)
    auto next_iteration = [&](int i, bool converged) {
        System& s = sys[i];
        OptimizerResult& r = out[i];
        if (converged || r.iterations >= opt.max_iterations) {
            s.phase = kDone;
            r.converged = converged;
            return;
        }
        ++r.iterations;
        s.d = lbfgs_direction(s.g, s.s_hist, s.y_hist, s.rho_hist);
        if (dot(s.d, s.g) >= 0.0 || !all_finite(s.d))
            for (size_t k = 0; k < s.d.size(); ++k) s.d[k] = -s.g[k];
        limit_atom_step(s.d, opt.max_atom_step);
        s.dd = dot(s.g, s.d);
        s.alpha = alpha0;
        s.tries = 0;
        s.last_err = kOk;
        s.last_msg.clear();
        set_trial(i);
    };

    // a warm start replaces the batch's start charges and start-charge model;
    // the caller's are back in place when the optimisation ends
    struct RestoreStart {
        Batch& b;
        const GuessModel* g;
        std::vector<double> q0;
        ~RestoreStart() {
            (void)b.set_guess(g);
            b.restore_start_charges(q0);
        }
    } restore{batch, batch.guess(), batch.start_charges()};
    std::vector<unsigned char> select(n, 0);
    std::vector<int> status;
    std::vector<double> energy, grad, qall;
    OptimizerStats st_local;
    while (true) {
        int pending = 0;
        for (int i = 0; i < n; ++i) {
            const System& s = sys[i];
            select[i] = (s.phase == kStart || s.phase == kTrial) ? 1 : 0;
            pending += select[i];
            const std::vector<double>& p = s.phase == kTrial ? s.trial : s.x;
            std::copy(p.begin(), p.end(), xall.begin() + 3 * (size_t)s.a0);
        }
        if (!pending) break;
        Status st = batch.set_positions(xall.data());
        if (!st.ok()) return st;
        st = batch.run(select.data());
        if (!st.ok()) return st;
        st = batch.fetch_energy_gradient(status, energy, grad, opt.warm_start ? &qall : nullptr);
        if (!st.ok()) return st;
        ++st_local.rounds;
        st_local.evaluations += pending;
        for (int i = 0; i < n; ++i) {
            if (!select[i]) continue;
            System& s = sys[i];
            OptimizerResult& r = out[i];
            const bool ok = status[i] == kOk;
            auto gslice = [&]() {
                return std::vector<double>(grad.begin() + 3 * (size_t)s.a0, grad.begin() + 3 * (size_t)(s.a0 + s.nat));
            };
            auto keep_charges = [&]() {
                if (!opt.warm_start) return;
                s.q.assign(qall.begin() + batch.shell_offset(i), qall.begin() + batch.shell_offset(i) + s.nsh);
            };
            if (s.phase == kStart) {
                // the start geometry must evaluate (src/optimizer.rs:93)
                if (!ok) {
                    s.phase = kFailed;
                    r.status = status[i];
                    r.message = batch.message(i);
                    continue;
                }
                s.e = energy[i];
                s.g = gslice();
                s.maxg = max_gradient(s.g);
                keep_charges();
                record(i, 0, 0.0);
                next_iteration(i, s.maxg <= opt.gradient_tolerance);
                continue;
            }
            // line-search trial (src/optimizer.rs:224-257)
            if (ok && energy[i] <= s.e + kArmijo * s.alpha * s.dd) {
                std::vector<double> g1 = gslice();
                std::vector<double> sv(s.x.size()), yv(s.x.size());
                for (size_t k = 0; k < sv.size(); ++k) sv[k] = s.trial[k] - s.x[k];
                for (size_t k = 0; k < yv.size(); ++k) yv[k] = g1[k] - s.g[k];
                const double step_norm = std::sqrt(dot(sv, sv));
                const double ys = dot(yv, sv);
                // history update (src/optimizer.rs:137-151)
                if (ys > 1.0e-12 && all_finite(sv) && all_finite(yv)) {
                    if (s.s_hist.size() == history_len) {
                        s.s_hist.erase(s.s_hist.begin());
                        s.y_hist.erase(s.y_hist.begin());
                        s.rho_hist.erase(s.rho_hist.begin());
                    }
                    s.s_hist.push_back(std::move(sv));
                    s.y_hist.push_back(std::move(yv));
                    s.rho_hist.push_back(1.0 / ys);
                }
                s.x = s.trial;
                s.e = energy[i];
                s.g = std::move(g1);
                s.maxg = max_gradient(s.g);
                keep_charges();
                if (opt.warm_start) {
                    st = batch.set_start_charges(i, s.q.data());
                    if (!st.ok()) return st;
                }
                record(i, r.iterations, step_norm);
                next_iteration(i, s.maxg <= opt.gradient_tolerance || step_norm <= opt.step_tolerance);
                continue;
            }
            if (!ok) {
                s.last_err = status[i];
                s.last_msg = batch.message(i);
            }
            s.alpha *= 0.5;
            if (++s.tries >= kLineSearchTrials) {
                // src/optimizer.rs:250-256
                s.phase = kFailed;
                r.status = s.last_err != kOk ? s.last_err : kErrInvalidInput;
                r.message = s.last_err != kOk ? s.last_msg : "L-BFGS line search failed to find a downhill step";
                continue;
            }
            set_trial(i);
        }
        // the first accepted start charges of a warm start replace the batch's guess
        if (opt.warm_start && st_local.rounds == 1) {
            st = batch.set_guess(nullptr);
            if (!st.ok()) return st;
            for (int i = 0; i < n; ++i)
                if (sys[i].phase != kFailed && !sys[i].q.empty()) {
                    st = batch.set_start_charges(i, sys[i].q.data());
                    if (!st.ok()) return st;
                }
        }
    }
    for (int i = 0; i < n; ++i) {
        OptimizerResult& r = out[i];
        if (r.status != kOk) continue;
        const System& s = sys[i];
        r.energy = s.e;
        r.max_gradient = s.maxg;
        r.positions = s.x;
        r.gradient = s.g;
    }
    st_local.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st_local;
    return Status{};
}

}  // namespace gfn1
