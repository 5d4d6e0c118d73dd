// Batched velocity-Verlet molecular dynamics (host driver): every step is one
// batched energy + gradient pass over the systems still running; the
// integration is host work on 3 nat vectors.  The reference runs dynamics
// through ASE (its calculator supplies energies and forces to ase.md).
#include "md.hpp"

#include <cmath>

namespace gfn1 {
namespace {

// IUPAC standard atomic weights, Z = 1..86 (the reference's table,
// data_tables.rs:40-128), in u
const double kMass[86] = {
    1.008,       4.002602,  6.94,      9.0121831, 10.81,     12.011,     14.007,    15.999,     18.998403163,
    20.1797,     22.98976928, 24.305,  26.9815385, 28.085,   30.973761998, 32.06,   35.45,      39.948,
    39.0983,     40.078,    44.955908, 47.867,    50.9415,   51.9961,    54.938044, 55.845,     58.933194,
    58.6934,     63.546,    65.38,     69.723,    72.630,    74.921595,  78.971,    79.904,     83.798,
    85.4678,     87.62,     88.90584,  91.224,    92.90637,  95.95,      98.0,      101.07,     102.90550,
    106.42,      107.8682,  112.414,   114.818,   118.710,   121.760,    127.60,    126.90447,  131.293,
    132.90545196, 137.327,  138.90547, 140.116,   140.90766, 144.242,    145.0,     150.36,     151.964,
    157.25,      158.92535, 162.500,   164.93033, 167.259,   168.93422,  173.045,   174.9668,   178.49,
    180.94788,   183.84,    186.207,   190.23,    192.217,   195.084,    196.966569, 200.592,   204.38,
    207.2,       208.98040, 209.0,     210.0,     222.0};
constexpr double kAmuToElectronMass = 1822.888486209;   // CODATA 2018
constexpr double kFsPerAtomicTime = 0.024188843265857;  // CODATA 2018

}  // namespace

Status md_batch(Batch& batch, const MdOptions& opt, std::vector<double>& v, std::vector<double>& pot,
                std::vector<double>& kin, std::vector<int>& status, std::vector<int>& iters) {
    const int n = batch.size(), na = batch.natoms();
    if (opt.steps < 0 || !(opt.timestep_fs > 0.0))
        return Status::error(kErrInvalidInput, "MD needs steps >= 0 and a positive time step");
    if ((int)v.size() != 3 * na) return Status::error(kErrInvalidInput, "one velocity per atom coordinate needed");
    std::vector<double> m(na), x(3 * (size_t)na), g, e, q;
    for (int a = 0; a < na; ++a) {
        const int z = batch.atom_number(a);
        m[a] = (z >= 1 && z <= 86 ? kMass[z - 1] : 1.0) * kAmuToElectronMass;
    }
    if (na > 0) {
        Status st = batch.positions(x.data());
        if (!st.ok()) return st;
    }
    const double dt = opt.timestep_fs / kFsPerAtomicTime;
    status.assign(n, kOk);
    pot.assign((size_t)(opt.steps + 1) * n, NAN);
    kin.assign((size_t)(opt.steps + 1) * n, NAN);
    iters.assign((size_t)(opt.steps + 1) * n, -1);
    // A warm start replaces the batch's start charges and, after the first
    // step, its learned start-charge model (whose prediction would otherwise
    // overwrite the previous step's charges); the caller's are put back when
    // the dynamics ends, on every exit path.
    struct RestoreStart {
        Batch& b;
        const GuessModel* g;
        std::vector<double> q0;
        ~RestoreStart() {
            (void)b.set_guess(g);
            b.restore_start_charges(q0);
        }
    } restore{batch, batch.guess(), batch.start_charges()};
    std::vector<unsigned char> select(n, 1);
    std::vector<int> st_run;
    auto evaluate = [&](int step) -> Status {
        for (int i = 0; i < n; ++i) select[i] = status[i] == kOk ? 1 : 0;
        Status st = batch.set_positions(x.data());
        if (!st.ok()) return st;
        st = batch.run(select.data());
        if (!st.ok()) return st;
        st = batch.fetch_energy_gradient(st_run, e, g, opt.warm_start ? &q : nullptr);
        if (!st.ok()) return st;
        for (int i = 0; i < n; ++i) {
            if (status[i] != kOk) continue;
            if (st_run[i] != kOk) {
                status[i] = st_run[i];
                continue;
            }
            if (opt.warm_start) {
                st = batch.set_start_charges(i, q.data() + batch.shell_offset(i));
                if (!st.ok()) return st;
            }
            double k = 0.0;
            for (int a = batch.atom_offset(i); a < batch.atom_offset(i + 1); ++a)
                for (int d = 0; d < 3; ++d) k += 0.5 * m[a] * v[3 * (size_t)a + d] * v[3 * (size_t)a + d];
            pot[(size_t)step * n + i] = e[i];
            kin[(size_t)step * n + i] = k;
            iters[(size_t)step * n + i] = batch.iterations(i);
        }
        return Status{};
    };
    // v <- v - dt/2 g / m for the running systems
    auto half_kick = [&]() {
        for (int i = 0; i < n; ++i) {
            if (status[i] != kOk) continue;
            for (size_t k = 3 * (size_t)batch.atom_offset(i); k < 3 * (size_t)batch.atom_offset(i + 1); ++k)
                v[k] -= 0.5 * dt * g[k] / m[k / 3];
        }
    };
    Status st = evaluate(0);
    if (!st.ok()) return st;
    if (opt.warm_start && batch.guess()) {
        st = batch.set_guess(nullptr);
        if (!st.ok()) return st;
    }
    for (int step = 1; step <= opt.steps; ++step) {
        half_kick();
        for (int i = 0; i < n; ++i) {
            if (status[i] != kOk) continue;
            for (size_t k = 3 * (size_t)batch.atom_offset(i); k < 3 * (size_t)batch.atom_offset(i + 1); ++k)
                x[k] += dt * v[k];
        }
        st = evaluate(step);
        if (!st.ok()) return st;
        half_kick();
        // the kinetic energy of this step is the one after the second half kick
        for (int i = 0; i < n; ++i) {
            if (status[i] != kOk) continue;
            double k = 0.0;
            for (int a = batch.atom_offset(i); a < batch.atom_offset(i + 1); ++a)
                for (int d = 0; d < 3; ++d) k += 0.5 * m[a] * v[3 * (size_t)a + d] * v[3 * (size_t)a + d];
            kin[(size_t)step * n + i] = k;
        }
    }
    return Status{};
}

}  // namespace gfn1
