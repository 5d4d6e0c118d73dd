// C ABI of the engine (include/gfn1_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gfn1_amd.h"
#include "engine.hpp"
#include "model.hpp"
#include "hessian.hpp"
#include "md.hpp"
#include "optimizer.hpp"

using namespace gfn1;

struct gfn1_params {
    Model model;
    std::mutex mu;
    std::vector<std::unique_ptr<DeviceModel>> per_device;  // lazily created per device id
    const DeviceModel* device_model(Status& st) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) {
            st = Status::error(kErrDevice, "no HIP device available");
            return nullptr;
        }
        std::lock_guard<std::mutex> lock(mu);
        if ((int)per_device.size() <= dev) per_device.resize(dev + 1);
        if (!per_device[dev]) {
            auto dm = std::make_unique<DeviceModel>();
            st = dm->init(model);
            if (!st.ok()) return nullptr;
            per_device[dev] = std::move(dm);
        }
        return per_device[dev].get();
    }
};

struct gfn1_electronic_result {
    InstanceResult r;
    double scalar[6] = {};
};

struct gfn1_gradient_result {
    gfn1_electronic_result electronic;
    std::vector<double> forces;
    double max_gradient = 0.0;
};

struct gfn1_guess_model {
    GuessModel m;
};

struct gfn1_xyz {
    std::vector<int32_t> z;
    std::vector<double> x;
    std::vector<double> lattice;  // 9 values when the input has a Lattice= field
    gfn1_system sys{};
};

struct gfn1_optimization {
    std::vector<OptimizerResult> results;
    OptimizerStats stats;
};

struct gfn1_batch {
    std::unique_ptr<Batch> batch;
    bool learned_guess = false;   // GFN1_GUESS_LEARNED requested
    bool guess_set = false;
    mutable std::vector<InstanceResult> results;
    mutable bool downloaded = false;
    bool ran = false;
    // results are copied to the host on the first accessor call after a run
    bool ready() const {
        if (!ran) return false;
        if (!downloaded) downloaded = batch->download(results).ok();
        return downloaded;
    }
};

namespace {

int fail(gfn1_error* err, const Status& st) {
    if (err) {
        err->code = st.code;
        err->iterations = 0;
        err->rms = 0.0;
        std::snprintf(err->message, sizeof(err->message), "%s", st.message.c_str());
    }
    return st.code;
}

void clear(gfn1_error* err) {
    if (err) {
        err->code = GFN1_OK;
        err->iterations = 0;
        err->rms = 0.0;
        err->message[0] = 0;
    }
}

EngineOptions to_engine(const gfn1_electronic_options& o) {
    EngineOptions e;
    e.max_scc = o.max_scc;
    e.energy_tolerance = o.energy_tolerance;
    e.charge_tolerance = o.charge_tolerance;
    e.mixing = o.mixing;
    e.scc_broyden = o.scc_broyden;
    e.scc_broyden_size = o.scc_broyden_size;
    e.level_shift = o.level_shift;
    e.eigen_tolerance = o.eigen_tolerance;
    e.electronic_temperature = o.electronic_temperature;
    e.enable_dispersion = o.enable_dispersion;
    e.integral_cutoff = o.integral_cutoff;
    e.coordination_cutoff = o.coordination_cutoff;
    e.enable_cn_hamiltonian = o.enable_cn_hamiltonian;
    return e;
}

Status validate(const gfn1_system* s, const gfn1_electronic_options* o) {
    if (!s || s->nat < 0 || (s->nat > 0 && (!s->numbers || !s->positions)))
        return Status::error(kErrInvalidInput, "invalid system");
    if (o && o->nprim != 0)
        return Status::error(kErrInvalidInput, "only the standard GFN1 primitive-count rule (nprim = 0) is supported");
    if (o && o->max_scc < 1) return Status::error(kErrInvalidInput, "max_scc must be at least 1");
    if (s->lattice)
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(s->lattice[k])) return Status::error(kErrInvalidInput, "non-finite lattice vector");
    return Status{};
}

SystemInput to_input(const gfn1_system* s, const gfn1_electronic_options* o) {
    SystemInput in;
    in.nat = s->nat;
    in.z = s->numbers;
    in.x = s->positions;
    in.charge = (o && o->has_charge) ? o->charge : s->charge;
    in.cell = s->lattice;
    if (o && o->scc_initial_shell_charges) {
        in.q0 = o->scc_initial_shell_charges;
        in.q0_len = o->scc_initial_shell_charges_len;
    }
    return in;
}

int run_one(const gfn1_system* system, const gfn1_params* params, const gfn1_electronic_options* eo,
            const EngineOptions& e, InstanceResult& out, gfn1_error* err) {
    clear(err);
    Status st = validate(system, eo);
    if (!st.ok()) return fail(err, st);
    if (!params) return fail(err, Status::error(kErrInvalidInput, "null parameters"));
    const DeviceModel* dm = const_cast<gfn1_params*>(params)->device_model(st);
    if (!dm) return fail(err, st);
    Batch b;
    std::vector<SystemInput> sys{to_input(system, eo)};
    st = b.create(*dm, sys, e);
    if (!st.ok()) return fail(err, st);
    st = b.run();
    if (!st.ok()) return fail(err, st);
    std::vector<InstanceResult> res;
    st = b.download(res);
    if (!st.ok()) return fail(err, st);
    out = std::move(res[0]);
    if (out.status != kOk) {
        Status s2 = Status::error(out.status, out.message.empty() ? "SCC did not converge" : out.message);
        int code = fail(err, s2);
        if (err && out.status == kErrSccNotConverged) {
            err->iterations = out.iterations;
            err->rms = out.rms;
        }
        return code;
    }
    return GFN1_OK;
}

void fill_scalars(gfn1_electronic_result* r) {
    r->scalar[GFN1_SCALAR_ELECTRONIC_TEMPERATURE] = r->r.temperature;
    r->scalar[GFN1_SCALAR_FERMI_LEVEL] = r->r.fermi_level;
    r->scalar[GFN1_SCALAR_NELEC] = r->r.nelec;
    r->scalar[GFN1_SCALAR_ITERATIONS] = r->r.iterations;
    r->scalar[GFN1_SCALAR_CONVERGED] = r->r.converged;
    r->scalar[GFN1_SCALAR_CHARGE_ORDER] = 3;
}

}  // namespace

extern "C" {

const char* gfn1_version(void) { return "gfn1-mi355x 0.1.0 (GFN1-xTB SCC energy + gradient, gfx950)"; }

int gfn1_params_load(const char* spec, gfn1_params** out, gfn1_error* err) {
    clear(err);
    if (!out) return fail(err, Status::error(kErrInvalidInput, "null output"));
    auto p = std::make_unique<gfn1_params>();
    Status st = Model::load(spec, p->model);
    if (!st.ok()) return fail(err, st);
    *out = p.release();
    return GFN1_OK;
}

int gfn1_params_from_text(const char* text, gfn1_params** out, gfn1_error* err) {
    clear(err);
    if (!out || !text) return fail(err, Status::error(kErrInvalidInput, "null argument"));
    auto p = std::make_unique<gfn1_params>();
    Status st = Model::load_text(text, p->model);
    if (!st.ok()) return fail(err, st);
    *out = p.release();
    return GFN1_OK;
}

void gfn1_params_free(gfn1_params* params) { delete params; }

void gfn1_electronic_options_default(gfn1_electronic_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->has_charge = 0;
    o->charge = 0.0;
    o->max_scc = 250;
    o->energy_tolerance = 1e-6;
    o->charge_tolerance = 2e-5;
    o->mixing = 0.4;
    o->scc_broyden = 1;
    o->scc_broyden_size = 250;
    o->level_shift = 0.0;
    o->nprim = 0;
    o->eigen_tolerance = 1e-12;
    o->electronic_temperature = 300.0;
    o->enable_dispersion = 1;
    o->integral_cutoff = 30.0;
    o->coordination_cutoff = 30.0;
    o->enable_cn_hamiltonian = 1;
    o->scc_initial_shell_charges = nullptr;
    o->scc_initial_shell_charges_len = 0;
}

void gfn1_gradient_options_default(gfn1_gradient_options* o) {
    gfn1_electronic_options_default(&o->electronic);
    o->include_repulsion = 1;
    o->include_dispersion = 1;
    o->include_hamiltonian = 1;
    o->include_scc = 1;
    o->include_halogen = 1;
}

int gfn1_run_electronic(const gfn1_system* system, const gfn1_params* params, const gfn1_electronic_options* options,
                        gfn1_electronic_result** out, gfn1_error* err) {
    if (!out) return fail(err, Status::error(kErrInvalidInput, "null output"));
    gfn1_electronic_options def;
    if (!options) {
        gfn1_electronic_options_default(&def);
        options = &def;
    }
    EngineOptions e = to_engine(*options);
    e.compute_gradient = 0;
    e.keep_matrices = 1;
    auto r = std::make_unique<gfn1_electronic_result>();
    int rc = run_one(system, params, options, e, r->r, err);
    if (rc != GFN1_OK) return rc;
    fill_scalars(r.get());
    *out = r.release();
    return GFN1_OK;
}

int gfn1_analytic_gradient(const gfn1_system* system, const gfn1_params* params,
                           const gfn1_gradient_options* options, gfn1_gradient_result** out, gfn1_error* err) {
    if (!out) return fail(err, Status::error(kErrInvalidInput, "null output"));
    gfn1_gradient_options def;
    if (!options) {
        gfn1_gradient_options_default(&def);
        options = &def;
    }
    EngineOptions e = to_engine(options->electronic);
    e.compute_gradient = 1;
    e.keep_matrices = 1;
    e.include_repulsion = options->include_repulsion;
    e.include_dispersion = options->include_dispersion;
    e.include_hamiltonian = options->include_hamiltonian;
    e.include_scc = options->include_scc;
    e.include_halogen = options->include_halogen;
    auto r = std::make_unique<gfn1_gradient_result>();
    int rc = run_one(system, params, &options->electronic, e, r->electronic.r, err);
    if (rc != GFN1_OK) return rc;
    fill_scalars(&r->electronic);
    const auto& g = r->electronic.r.gradient;
    r->forces.resize(g.size());
    r->max_gradient = 0.0;
    for (size_t k = 0; k < g.size(); ++k) {
        r->forces[k] = -g[k];
        r->max_gradient = std::max(r->max_gradient, std::fabs(g[k]));
    }
    *out = r.release();
    return GFN1_OK;
}

void gfn1_electronic_result_free(gfn1_electronic_result* r) { delete r; }
void gfn1_gradient_result_free(gfn1_gradient_result* r) { delete r; }

void gfn1_electronic_result_sizes(const gfn1_electronic_result* r, int32_t* nat, int32_t* nsh, int32_t* nao) {
    if (nat) *nat = r->r.nat;
    if (nsh) *nsh = r->r.nsh;
    if (nao) *nao = r->r.nao;
}

void gfn1_electronic_result_energy_terms(const gfn1_electronic_result* r, double out[15]) {
    std::memcpy(out, r->r.energies, sizeof(r->r.energies));
}

const double* gfn1_electronic_result_matrix(const gfn1_electronic_result* r, int field) {
    switch (field) {
        case GFN1_MAT_OVERLAP: return r->r.S.data();
        case GFN1_MAT_H0: return r->r.H0.data();
        case GFN1_MAT_FOCK: return r->r.F.data();
        case GFN1_MAT_DENSITY: return r->r.P.data();
        case GFN1_MAT_ENERGY_WEIGHTED_DENSITY: return r->r.W.data();
        default: return nullptr;
    }
}

const double* gfn1_electronic_result_vector(const gfn1_electronic_result* r, int field) {
    switch (field) {
        case GFN1_VEC_ORBITAL_ENERGIES: return r->r.eps.data();
        case GFN1_VEC_OCCUPATIONS: return r->r.occ.data();
        case GFN1_VEC_SHELL_CHARGES: return r->r.shell_charges.data();
        case GFN1_VEC_ATOMIC_CHARGES: return r->r.atomic_charges.data();
        case GFN1_VEC_SHELL_SCC_POTENTIAL: return r->r.shell_potential.data();
        case GFN1_VEC_COORDINATION_NUMBERS: return r->r.cn.data();
        case GFN1_VEC_DIPOLE: return r->r.dipole;
        default: return nullptr;
    }
}

double gfn1_electronic_result_scalar(const gfn1_electronic_result* r, int field) {
    if (field < 0 || field > 5) return NAN;
    return r->scalar[field];
}

const gfn1_electronic_result* gfn1_gradient_result_electronic(const gfn1_gradient_result* r) { return &r->electronic; }
double gfn1_gradient_result_total_energy(const gfn1_gradient_result* r) { return r->electronic.r.energies[0]; }
double gfn1_gradient_result_max_gradient(const gfn1_gradient_result* r) { return r->max_gradient; }

const double* gfn1_gradient_result_field(const gfn1_gradient_result* r, int field) {
    const InstanceResult& e = r->electronic.r;
    switch (field) {
        case GFN1_GRAD_TOTAL: return e.gradient.data();
        case GFN1_GRAD_FORCES: return r->forces.data();
        case GFN1_GRAD_ELECTRONIC: return e.g_el.data();
        case GFN1_GRAD_REPULSION: return e.g_rep.data();
        case GFN1_GRAD_DISPERSION: return e.g_disp.data();
        case GFN1_GRAD_HALOGEN: return e.g_hal.data();
        default: return nullptr;
    }
}

void gfn1_batch_options_default(gfn1_batch_options* o) {
    gfn1_gradient_options_default(&o->gradient);
    o->precision = GFN1_PRECISION_FP64;
    o->guess = GFN1_GUESS_ZERO;
    o->eigensolver = 0;
    o->keep_matrices = 0;
    o->scc_graphs = 1;
}

int gfn1_batch_create(int32_t n_inst, const gfn1_system* systems, const gfn1_params* params,
                      const gfn1_batch_options* options, gfn1_batch** out, gfn1_error* err) {
    clear(err);
    if (!out || n_inst < 0 || (n_inst > 0 && !systems) || !params)
        return fail(err, Status::error(kErrInvalidInput, "invalid batch arguments"));
    gfn1_batch_options def;
    if (!options) {
        gfn1_batch_options_default(&def);
        options = &def;
    }
    if (options->precision != GFN1_PRECISION_FP64 && options->precision != GFN1_PRECISION_MIXED)
        return fail(err, Status::error(kErrInvalidInput, "unknown precision mode"));
    if (options->guess != GFN1_GUESS_ZERO && options->guess != GFN1_GUESS_LEARNED)
        return fail(err, Status::error(kErrInvalidInput, "unknown charge guess mode"));
    std::vector<SystemInput> sys;
    for (int i = 0; i < n_inst; ++i) {
        Status st = validate(&systems[i], &options->gradient.electronic);
        if (!st.ok()) return fail(err, st);
        sys.push_back(to_input(&systems[i], &options->gradient.electronic));
    }
    EngineOptions e = to_engine(options->gradient.electronic);
    e.compute_gradient = 1;
    e.include_repulsion = options->gradient.include_repulsion;
    e.include_dispersion = options->gradient.include_dispersion;
    e.include_hamiltonian = options->gradient.include_hamiltonian;
    e.include_scc = options->gradient.include_scc;
    e.include_halogen = options->gradient.include_halogen;
    e.eigensolver = options->eigensolver;
    e.keep_matrices = options->keep_matrices;
    e.scc_graphs = options->scc_graphs;
    e.mixed = options->precision == GFN1_PRECISION_MIXED;
    Status st;
    const DeviceModel* dm = const_cast<gfn1_params*>(params)->device_model(st);
    if (!dm) return fail(err, st);
    auto b = std::make_unique<gfn1_batch>();
    b->batch = std::make_unique<Batch>();
    st = b->batch->create(*dm, sys, e);
    if (!st.ok()) return fail(err, st);
    b->learned_guess = options->guess == GFN1_GUESS_LEARNED;
    *out = b.release();
    return GFN1_OK;
}

int gfn1_batch_set_guess(gfn1_batch* b, const gfn1_guess_model* model, gfn1_error* err) {
    clear(err);
    if (!b || !b->batch) return fail(err, Status::error(kErrInvalidInput, "invalid batch"));
    Status st = b->batch->set_guess(model ? &model->m : nullptr);
    if (!st.ok()) return fail(err, st);
    b->guess_set = model != nullptr;
    b->learned_guess = model != nullptr;
    return GFN1_OK;
}

int gfn1_batch_run(gfn1_batch* batch, gfn1_error* err) {
    clear(err);
    if (batch && batch->learned_guess && !batch->guess_set)
        return fail(err, Status::error(kErrInvalidInput, "learned charge guess selected but no model set"));
    if (!batch) return fail(err, Status::error(kErrInvalidInput, "null batch"));
    batch->downloaded = false;
    Status st = batch->batch->run();
    batch->ran = st.ok();
    if (!st.ok()) return fail(err, st);
    return GFN1_OK;
}

int gfn1_batch_status(const gfn1_batch* b, int32_t inst) {
    if (!b || !b->ready() || inst < 0 || inst >= (int)b->results.size()) return GFN1_ERR_INVALID_INPUT;
    return b->results[inst].status;
}

double gfn1_batch_energy(const gfn1_batch* b, int32_t inst) {
    if (!b || !b->ready() || inst < 0 || inst >= (int)b->results.size()) return NAN;
    return b->results[inst].energies[0];
}

void gfn1_batch_energy_terms(const gfn1_batch* b, int32_t inst, double out[15]) {
    if (!b || !b->ready() || inst < 0 || inst >= (int)b->results.size()) return;
    std::memcpy(out, b->results[inst].energies, sizeof(double) * 15);
}

int32_t gfn1_batch_iterations(const gfn1_batch* b, int32_t inst) {
    if (!b || !b->ready() || inst < 0 || inst >= (int)b->results.size()) return -1;
    return b->results[inst].iterations;
}

int gfn1_batch_gradient(const gfn1_batch* b, int32_t inst, double* out) {
    if (!b || !b->ready() || inst < 0 || inst >= (int)b->results.size() || !out) return GFN1_ERR_INVALID_INPUT;
    const auto& g = b->results[inst].gradient;
    std::copy(g.begin(), g.end(), out);
    return b->results[inst].status;
}

int gfn1_batch_shell_charges(const gfn1_batch* b, int32_t inst, double* out) {
    if (!b || !b->ready() || inst < 0 || inst >= (int)b->results.size() || !out) return GFN1_ERR_INVALID_INPUT;
    const auto& q = b->results[inst].shell_charges;
    std::copy(q.begin(), q.end(), out);
    return b->results[inst].status;
}

int gfn1_batch_kernel_stats(const gfn1_batch* b, int kernel, double* total_ms, int64_t* launches, double* bytes) {
    if (!b || !b->batch || kernel != GFN1_KERNEL_TRIDIAG || !total_ms || !launches || !bytes) return GFN1_ERR_INVALID_INPUT;
    long long nl = 0;
    b->batch->tridiag_stats(total_ms, &nl, bytes);
    *launches = nl;
    return GFN1_OK;
}

int gfn1_batch_autotune(gfn1_batch* b, gfn1_error* err) {
    clear(err);
    if (!b || !b->batch) return fail(err, Status::error(kErrInvalidInput, "invalid batch"));
    Status st = b->batch->autotune();
    if (!st.ok()) return fail(err, st);
    return GFN1_OK;
}

int gfn1_batch_set_backtransform_width(gfn1_batch* b, int32_t cols, gfn1_error* err) {
    clear(err);
    if (!b || !b->batch) return fail(err, Status::error(kErrInvalidInput, "invalid batch"));
    Status st = b->batch->set_backtransform_width(cols);
    if (!st.ok()) return fail(err, st);
    return GFN1_OK;
}

int32_t gfn1_batch_tuning_count(const gfn1_batch* b) {
    return (b && b->batch) ? (int32_t)b->batch->tuning().size() : 0;
}

int gfn1_batch_tuning_entry(const gfn1_batch* b, int32_t k, int32_t out_i[3], double out_ms[3]) {
    if (!b || !b->batch || !out_i || !out_ms || k < 0 || k >= (int32_t)b->batch->tuning().size())
        return GFN1_ERR_INVALID_INPUT;
    const auto& r = b->batch->tuning()[k];
    out_i[0] = r.ld;
    out_i[1] = r.count;
    out_i[2] = r.cols;
    for (int w = 0; w < 3; ++w) out_ms[w] = r.ms[w];
    return GFN1_OK;
}

int gfn1_batch_launch_stats(const gfn1_batch* b, int64_t out[GFN1_LAUNCH_STATS]) {
    if (!b || !b->batch || !out) return GFN1_ERR_INVALID_INPUT;
    const LaunchStats& ls = b->batch->launch_stats();
    out[GFN1_LAUNCH_KERNELS] = ls.kernels;
    out[GFN1_LAUNCH_ENQUEUES] = ls.enqueues;
    out[GFN1_LAUNCH_GRAPH_REPLAYS] = ls.graph_replays;
    out[GFN1_LAUNCH_HOST_ITERATIONS] = ls.host_iterations;
    out[GFN1_LAUNCH_STATE_READS] = ls.state_reads;
    out[GFN1_LAUNCH_GRAPHS_CAPTURED] = ls.graphs_captured;
    out[GFN1_LAUNCH_CAPTURE_US] = (int64_t)(ls.capture_ms * 1000.0 + 0.5);
    return GFN1_OK;
}

void gfn1_batch_timings(const gfn1_batch* b, double out[4]) {
    if (!b) return;
    std::memcpy(out, b->batch->timings(), sizeof(double) * 4);
}

void gfn1_batch_free(gfn1_batch* b) { delete b; }

int gfn1_batch_set_positions(gfn1_batch* b, const double* positions, gfn1_error* err) {
    clear(err);
    if (!b || !b->batch) return fail(err, Status::error(kErrInvalidInput, "invalid batch"));
    Status st = b->batch->set_positions(positions);
    if (!st.ok()) return fail(err, st);
    b->ran = false;  // the results belong to the old positions
    b->downloaded = false;
    return GFN1_OK;
}

int gfn1_batch_statuses(const gfn1_batch* b, int32_t* out) {
    if (!b || !b->ready() || !out) return GFN1_ERR_INVALID_INPUT;
    for (size_t i = 0; i < b->results.size(); ++i) out[i] = b->results[i].status;
    return GFN1_OK;
}

int gfn1_batch_energies(const gfn1_batch* b, double* out) {
    if (!b || !b->ready() || !out) return GFN1_ERR_INVALID_INPUT;
    for (size_t i = 0; i < b->results.size(); ++i)
        out[i] = b->results[i].status == kOk ? b->results[i].energies[0] : NAN;
    return GFN1_OK;
}

int gfn1_batch_gradients(const gfn1_batch* b, double* out) {
    if (!b || !b->ready() || !out) return GFN1_ERR_INVALID_INPUT;
    for (const InstanceResult& r : b->results) {
        if (r.status == kOk) std::copy(r.gradient.begin(), r.gradient.end(), out);
        else std::fill(out, out + 3 * (size_t)r.nat, NAN);
        out += 3 * (size_t)r.nat;
    }
    return GFN1_OK;
}

int gfn1_batch_all_shell_charges(const gfn1_batch* b, double* out) {
    if (!b || !b->ready() || !out) return GFN1_ERR_INVALID_INPUT;
    for (size_t i = 0; i < b->results.size(); ++i) {
        const InstanceResult& r = b->results[i];
        const size_t nsh = (size_t)b->batch->shell_count((int)i);
        if (r.status == kOk) std::copy(r.shell_charges.begin(), r.shell_charges.end(), out);
        else std::fill(out, out + nsh, NAN);
        out += nsh;
    }
    return GFN1_OK;
}

int gfn1_batch_set_all_start_charges(gfn1_batch* b, const double* q, gfn1_error* err) {
    clear(err);
    if (!b || !b->batch || !q) return fail(err, Status::error(kErrInvalidInput, "invalid start charges"));
    for (int i = 0; i < b->batch->size(); ++i) {
        Status st = b->batch->set_start_charges(i, q);
        if (!st.ok()) return fail(err, st);
        q += b->batch->shell_count(i);
    }
    return GFN1_OK;
}

int gfn1_batch_set_start_charges(gfn1_batch* b, int32_t inst, const double* q, gfn1_error* err) {
    clear(err);
    if (!b || !b->batch || inst < 0 || inst >= b->batch->size())
        return fail(err, Status::error(kErrInvalidInput, "invalid batch instance"));
    Status st = b->batch->set_start_charges(inst, q);
    if (!st.ok()) return fail(err, st);
    return GFN1_OK;
}

int gfn1_batch_positions(const gfn1_batch* b, double* out) {
    if (!b || !b->batch || (!out && b->batch->natoms() > 0)) return GFN1_ERR_INVALID_INPUT;
    if (b->batch->natoms() > 0) b->batch->positions(out);
    return GFN1_OK;
}

void gfn1_md_options_default(gfn1_md_options* o) {
    if (!o) return;
    MdOptions d;
    o->timestep_fs = d.timestep_fs;
    o->steps = d.steps;
    o->warm_start = d.warm_start;
}

int gfn1_batch_md(gfn1_batch* b, const gfn1_md_options* options, double* velocities, double* potential,
                  double* kinetic, int32_t* status, int32_t* scc_iterations, gfn1_error* err) {
    clear(err);
    if (!b || !b->batch || (!velocities && b->batch->natoms() > 0))
        return fail(err, Status::error(kErrInvalidInput, "invalid MD arguments"));
    if (b->learned_guess && !b->guess_set)
        return fail(err, Status::error(kErrInvalidInput, "learned charge guess selected but no model set"));
    gfn1_md_options def;
    if (!options) {
        gfn1_md_options_default(&def);
        options = &def;
    }
    MdOptions o;
    o.timestep_fs = options->timestep_fs;
    o.steps = options->steps;
    o.warm_start = options->warm_start;
    const size_t nv = 3 * (size_t)b->batch->natoms();
    std::vector<double> v(velocities, velocities + nv), pot, kin;
    std::vector<int> st_out, it_out;
    b->ran = false;
    b->downloaded = false;
    Status st = md_batch(*b->batch, o, v, pot, kin, st_out, it_out);
    if (!st.ok()) return fail(err, st);
    std::copy(v.begin(), v.end(), velocities);
    if (potential) std::copy(pot.begin(), pot.end(), potential);
    if (kinetic) std::copy(kin.begin(), kin.end(), kinetic);
    if (status) std::copy(st_out.begin(), st_out.end(), status);
    if (scc_iterations) std::copy(it_out.begin(), it_out.end(), scc_iterations);
    return GFN1_OK;
}

void gfn1_optimization_options_default(gfn1_optimization_options* o) {
    if (!o) return;
    OptimizerOptions d;
    o->max_iterations = d.max_iterations;
    o->gradient_tolerance = d.gradient_tolerance;
    o->step_tolerance = d.step_tolerance;
    o->history = d.history;
    o->initial_step = d.initial_step;
    o->max_atom_step = d.max_atom_step;
    o->warm_start = d.warm_start;
    o->keep_trajectory = d.keep_trajectory;
}

int gfn1_batch_optimize(gfn1_batch* b, const gfn1_optimization_options* options, gfn1_optimization** out,
                        gfn1_error* err) {
    clear(err);
    if (!b || !b->batch || !out) return fail(err, Status::error(kErrInvalidInput, "invalid optimisation arguments"));
    if (b->learned_guess && !b->guess_set)
        return fail(err, Status::error(kErrInvalidInput, "learned charge guess selected but no model set"));
    gfn1_optimization_options def;
    if (!options) {
        gfn1_optimization_options_default(&def);
        options = &def;
    }
    OptimizerOptions o;
    o.max_iterations = options->max_iterations;
    o.gradient_tolerance = options->gradient_tolerance;
    o.step_tolerance = options->step_tolerance;
    o.history = options->history;
    o.initial_step = options->initial_step;
    o.max_atom_step = options->max_atom_step;
    o.warm_start = options->warm_start;
    o.keep_trajectory = options->keep_trajectory;
    auto r = std::make_unique<gfn1_optimization>();
    b->ran = false;
    b->downloaded = false;
    Status st = optimize_batch(*b->batch, o, r->results, &r->stats);
    if (!st.ok()) return fail(err, st);
    *out = r.release();
    return GFN1_OK;
}

int gfn1_optimize_geometry(const gfn1_system* system, const gfn1_params* params,
                           const gfn1_gradient_options* gradient_options, const gfn1_optimization_options* options,
                           gfn1_optimization** out, gfn1_error* err) {
    clear(err);
    if (!system || !out) return fail(err, Status::error(kErrInvalidInput, "invalid optimisation arguments"));
    if (system->nat == 0) return fail(err, Status::error(kErrInvalidInput, "cannot optimize an empty system"));
    gfn1_batch_options bo;
    gfn1_batch_options_default(&bo);
    if (gradient_options) bo.gradient = *gradient_options;
    gfn1_batch* b = nullptr;
    int rc = gfn1_batch_create(1, system, params, &bo, &b, err);
    if (rc != GFN1_OK) return rc;
    std::unique_ptr<gfn1_batch> hold(b);
    gfn1_optimization* r = nullptr;
    rc = gfn1_batch_optimize(b, options, &r, err);
    if (rc != GFN1_OK) return rc;
    const OptimizerResult& res = r->results[0];
    if (res.status != kOk) {
        const int code = res.status;
        fail(err, Status::error(code, res.message));
        gfn1_optimization_free(r);
        return code;
    }
    *out = r;
    return GFN1_OK;
}

namespace {
const OptimizerResult* opt_result(const gfn1_optimization* r, int32_t inst) {
    if (!r || inst < 0 || inst >= (int)r->results.size()) return nullptr;
    return &r->results[inst];
}
}  // namespace

int32_t gfn1_optimization_count(const gfn1_optimization* r) { return r ? (int32_t)r->results.size() : 0; }

int gfn1_optimization_status(const gfn1_optimization* r, int32_t inst) {
    const OptimizerResult* x = opt_result(r, inst);
    return x ? x->status : GFN1_ERR_INVALID_INPUT;
}

const char* gfn1_optimization_message(const gfn1_optimization* r, int32_t inst) {
    const OptimizerResult* x = opt_result(r, inst);
    return x ? x->message.c_str() : "";
}

double gfn1_optimization_energy(const gfn1_optimization* r, int32_t inst) {
    const OptimizerResult* x = opt_result(r, inst);
    return x && x->status == kOk ? x->energy : NAN;
}

double gfn1_optimization_max_gradient(const gfn1_optimization* r, int32_t inst) {
    const OptimizerResult* x = opt_result(r, inst);
    return x && x->status == kOk ? x->max_gradient : NAN;
}

int32_t gfn1_optimization_iterations(const gfn1_optimization* r, int32_t inst) {
    const OptimizerResult* x = opt_result(r, inst);
    return x ? x->iterations : -1;
}

int32_t gfn1_optimization_converged(const gfn1_optimization* r, int32_t inst) {
    const OptimizerResult* x = opt_result(r, inst);
    return x && x->converged ? 1 : 0;
}

int gfn1_optimization_positions(const gfn1_optimization* r, int32_t inst, double* out) {
    const OptimizerResult* x = opt_result(r, inst);
    if (!x || !out) return GFN1_ERR_INVALID_INPUT;
    std::copy(x->positions.begin(), x->positions.end(), out);
    return x->status;
}

int gfn1_optimization_gradient(const gfn1_optimization* r, int32_t inst, double* out) {
    const OptimizerResult* x = opt_result(r, inst);
    if (!x || !out) return GFN1_ERR_INVALID_INPUT;
    std::copy(x->gradient.begin(), x->gradient.end(), out);
    return x->status;
}

int32_t gfn1_optimization_trajectory_length(const gfn1_optimization* r, int32_t inst) {
    const OptimizerResult* x = opt_result(r, inst);
    return x ? (int32_t)x->trajectory.size() : 0;
}

int gfn1_optimization_trajectory_step(const gfn1_optimization* r, int32_t inst, int32_t step, int32_t* iteration,
                                      double* energy, double* max_gradient, double* step_norm, double* positions) {
    const OptimizerResult* x = opt_result(r, inst);
    if (!x || step < 0 || step >= (int32_t)x->trajectory.size()) return GFN1_ERR_INVALID_INPUT;
    const OptimizerStep& s = x->trajectory[step];
    if (iteration) *iteration = s.iteration;
    if (energy) *energy = s.energy;
    if (max_gradient) *max_gradient = s.max_gradient;
    if (step_norm) *step_norm = s.step_norm;
    if (positions) std::copy(s.positions.begin(), s.positions.end(), positions);
    return GFN1_OK;
}

void gfn1_optimization_stats(const gfn1_optimization* r, int32_t* rounds, int64_t* evaluations, double* total_ms) {
    if (!r) return;
    if (rounds) *rounds = r->stats.rounds;
    if (evaluations) *evaluations = r->stats.evaluations;
    if (total_ms) *total_ms = r->stats.total_ms;
}

void gfn1_optimization_free(gfn1_optimization* r) { delete r; }

namespace {
// Element symbols Z = 1..86 (src/system.rs:255-262).
const char* const kSymbols[] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar",
    "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr",
    "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe",
    "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf",
    "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn"};
// symbol_to_z (src/system.rs:154-165): an atomic number 1..86, or a symbol
// with its first letter upper-cased and the rest lower-cased
int symbol_to_z(const std::string& t) {
    if (!t.empty() && t.size() <= 3 && std::all_of(t.begin(), t.end(), [](char c) { return c >= '0' && c <= '9'; })) {
        const int z = std::atoi(t.c_str());
        if (z >= 1 && z <= 86) return z;
    }
    std::string s = t;
    for (size_t i = 0; i < s.size(); ++i) s[i] = i == 0 ? (char)std::toupper((unsigned char)s[i]) : (char)std::tolower((unsigned char)s[i]);
    for (int z = 1; z <= 86; ++z)
        if (s == kSymbols[z - 1]) return z;
    return 0;
}

bool parse_number(const std::string& t, double* v) {
    if (t.empty()) return false;
    char* end = nullptr;
    *v = std::strtod(t.c_str(), &end);
    return end == t.c_str() + t.size();
}

std::vector<std::string> split_ws(const std::string& line) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
        const size_t j = i;
        while (i < line.size() && !std::isspace((unsigned char)line[i])) ++i;
        if (i > j) out.push_back(line.substr(j, i - j));
    }
    return out;
}
}  // namespace

int gfn1_system_from_xyz(const char* text, double charge, int32_t coordinates_are_bohr, gfn1_xyz** out,
                         gfn1_error* err) {
    clear(err);
    if (!text || !out) return fail(err, Status::error(kErrInvalidInput, "invalid XYZ arguments"));
    std::vector<std::string> lines;
    {
        std::string cur;
        for (const char* p = text; *p; ++p) {
            if (*p == '\n') {
                if (!cur.empty() && cur.back() == '\r') cur.pop_back();
                lines.push_back(cur);
                cur.clear();
            } else {
                cur.push_back(*p);
            }
        }
        if (!cur.empty()) lines.push_back(cur);
    }
    if (lines.empty()) return fail(err, Status::error(kErrInvalidInput, "empty XYZ"));
    const std::vector<std::string> head = split_ws(lines[0]);
    char* end = nullptr;
    const long long natoms = head.size() == 1 ? std::strtoll(head[0].c_str(), &end, 10) : -1;
    if (head.size() != 1 || end != head[0].c_str() + head[0].size() || natoms < 0)
        return fail(err, Status::error(kErrInvalidInput, "invalid XYZ atom count: " + lines[0]));
    const std::string comment = lines.size() > 1 ? lines[1] : std::string();
    auto r = std::make_unique<gfn1_xyz>();
    const double scale = coordinates_are_bohr ? 1.0 : gfn1::kAngstromToBohr;
    // extended-XYZ cell (src/system.rs:191-254)
    const size_t lp = comment.find("Lattice=\"");
    if (lp != std::string::npos) {
        const size_t b0 = lp + 9, e0 = comment.find('"', b0);
        if (e0 == std::string::npos)
            return fail(err, Status::error(kErrInvalidInput, "unterminated Lattice=\"...\" field in XYZ comment"));
        const std::vector<std::string> vals = split_ws(comment.substr(b0, e0 - b0));
        for (const std::string& v : vals) {
            double d = 0.0;
            if (!parse_number(v, &d))
                return fail(err, Status::error(kErrInvalidInput, "invalid number in XYZ Lattice field"));
            r->lattice.push_back(d * scale);
        }
        if (r->lattice.size() != 9)
            return fail(err, Status::error(kErrInvalidInput, "XYZ Lattice field needs 9 numbers, got " +
                                                                 std::to_string(r->lattice.size())));
        size_t pp = comment.find("pbc=\"");
        if (pp == std::string::npos) pp = comment.find("PBC=\"");
        if (pp != std::string::npos) {
            const size_t b1 = pp + 5, e1 = comment.find('"', b1);
            if (e1 == std::string::npos)
                return fail(err, Status::error(kErrInvalidInput, "unterminated pbc=\"...\" field in XYZ comment"));
            const std::vector<std::string> toks = split_ws(comment.substr(b1, e1 - b1));
            if (toks.size() != 3)
                return fail(err, Status::error(kErrInvalidInput, "XYZ pbc field needs 3 boolean tokens, got " +
                                                                     std::to_string(toks.size())));
            for (const std::string& tk : toks) {
                std::string l = tk;
                for (char& ch : l) ch = (char)std::tolower((unsigned char)ch);
                const bool yes = l == "t" || l == "true" || l == "1" || l == ".true." || l == "yes" || l == "y";
                const bool no = l == "f" || l == "false" || l == "0" || l == ".false." || l == "no" || l == "n";
                if (!yes && !no) return fail(err, Status::error(kErrInvalidInput, "invalid XYZ pbc token: " + tk));
                if (no)
                    return fail(err, Status::error(kErrInvalidInput,
                                                   "only fully periodic cells (pbc=\"T T T\") are supported"));
            }
        }
        CellRec probe;
        std::vector<double> gprobe;
        if (!make_cell_record(r->lattice.data(), probe, gprobe))
            return fail(err, Status::error(kErrSingularCell, "singular cell"));
    }
    for (long long idx = 0; idx < natoms; ++idx) {
        const size_t line_no = (size_t)idx + 3;
        if ((size_t)idx + 2 >= lines.size())
            return fail(err, Status::error(kErrInvalidInput, "XYZ ended before atom " + std::to_string(idx)));
        const std::vector<std::string> parts = split_ws(lines[(size_t)idx + 2]);
        if (parts.size() < 4)
            return fail(err, Status::error(kErrInvalidInput,
                                           "XYZ line " + std::to_string(line_no) + " has fewer than 4 fields"));
        const int z = symbol_to_z(parts[0]);
        if (!z)
            return fail(err, Status::error(kErrInvalidInput, "unknown element on line " + std::to_string(line_no) +
                                                                 ": " + parts[0]));
        r->z.push_back(z);
        for (int k = 1; k <= 3; ++k) {
            double v = 0.0;
            if (!parse_number(parts[k], &v))
                return fail(err, Status::error(kErrParse, "line " + std::to_string(line_no) +
                                                              ": invalid floating point value: " + parts[k]));
            r->x.push_back(v * scale);
        }
    }
    r->sys.nat = (int32_t)r->z.size();
    r->sys.numbers = r->z.empty() ? nullptr : r->z.data();
    r->sys.positions = r->x.empty() ? nullptr : r->x.data();
    r->sys.charge = charge;
    r->sys.lattice = r->lattice.empty() ? nullptr : r->lattice.data();
    *out = r.release();
    return GFN1_OK;
}

const gfn1_system* gfn1_xyz_system(const gfn1_xyz* x) { return x ? &x->sys : nullptr; }

void gfn1_xyz_free(gfn1_xyz* x) { delete x; }

int32_t gfn1_params_source(const gfn1_params* params, char* out, int32_t len) {
    if (!params) return -1;
    const std::string d = params->model.source_description();
    if (out && len > 0) {
        const size_t k = std::min((size_t)(len - 1), d.size());
        std::memcpy(out, d.data(), k);
        out[k] = 0;
    }
    return (int32_t)d.size();
}

void gfn1_hessian_options_default(gfn1_hessian_options* o) {
    if (!o) return;
    HessianOptions d;
    o->step = d.step;
    o->scc_energy_tolerance = d.scc_energy_tolerance;
    o->scc_charge_tolerance = d.scc_charge_tolerance;
    o->include_repulsion = d.include_repulsion;
    o->include_electronic = d.include_electronic;
    o->include_dispersion = d.include_dispersion;
    o->include_halogen = d.include_halogen;
    o->warm_start = d.warm_start;
    o->max_batch = d.max_batch;
}

namespace {
int hessian_common(int32_t n, const gfn1_system* systems, const gfn1_params* params,
                   const gfn1_electronic_options* eo, const gfn1_hessian_options* options,
                   std::vector<HessianResult>& res, gfn1_error* err) {
    clear(err);
    if (n < 0 || (n > 0 && !systems) || !params)
        return fail(err, Status::error(kErrInvalidInput, "invalid Hessian arguments"));
    gfn1_electronic_options edef;
    if (!eo) {
        gfn1_electronic_options_default(&edef);
        eo = &edef;
    }
    gfn1_hessian_options hdef;
    if (!options) {
        gfn1_hessian_options_default(&hdef);
        options = &hdef;
    }
    std::vector<SystemInput> sys;
    for (int i = 0; i < n; ++i) {
        Status st = validate(&systems[i], eo);
        if (!st.ok()) return fail(err, st);
        sys.push_back(to_input(&systems[i], eo));
    }
    HessianOptions o;
    o.step = options->step;
    o.scc_energy_tolerance = options->scc_energy_tolerance;
    o.scc_charge_tolerance = options->scc_charge_tolerance;
    o.include_repulsion = options->include_repulsion;
    o.include_electronic = options->include_electronic;
    o.include_dispersion = options->include_dispersion;
    o.include_halogen = options->include_halogen;
    o.warm_start = options->warm_start;
    o.max_batch = options->max_batch;
    Status st;
    const DeviceModel* dm = const_cast<gfn1_params*>(params)->device_model(st);
    if (!dm) return fail(err, st);
    st = hessian_batch(*dm, sys, to_engine(*eo), o, res, nullptr);
    if (!st.ok()) return fail(err, st);
    return GFN1_OK;
}
}  // namespace

int gfn1_hessian(const gfn1_system* system, const gfn1_params* params, const gfn1_electronic_options* electronic,
                 const gfn1_hessian_options* options, double* hessian, double* max_asymmetry,
                 double* dipole_derivatives, double* dipole, gfn1_error* err) {
    if (!system || !hessian) {
        clear(err);
        return fail(err, Status::error(kErrInvalidInput, "invalid Hessian arguments"));
    }
    std::vector<HessianResult> res;
    int rc = hessian_common(1, system, params, electronic, options, res, err);
    if (rc != GFN1_OK) return rc;
    const HessianResult& r = res[0];
    if (r.status != kOk) return fail(err, Status::error(r.status, r.message));
    std::copy(r.hessian.begin(), r.hessian.end(), hessian);
    if (max_asymmetry) *max_asymmetry = r.max_asymmetry;
    if (dipole_derivatives) std::copy(r.dipole_derivatives.begin(), r.dipole_derivatives.end(), dipole_derivatives);
    if (dipole) std::copy(r.dipole, r.dipole + 3, dipole);
    return GFN1_OK;
}

int gfn1_hessian_terms(const gfn1_system* system, const gfn1_params* params, const gfn1_electronic_options* electronic,
                       const gfn1_hessian_options* options, double* hessian, double* max_asymmetry,
                       double* repulsion, double* dispersion, double* halogen, gfn1_error* err) {
    if (!system || !hessian) {
        clear(err);
        return fail(err, Status::error(kErrInvalidInput, "invalid Hessian arguments"));
    }
    std::vector<HessianResult> res;
    int rc = hessian_common(1, system, params, electronic, options, res, err);
    if (rc != GFN1_OK) return rc;
    const HessianResult& r = res[0];
    if (r.status != kOk) return fail(err, Status::error(r.status, r.message));
    std::copy(r.hessian.begin(), r.hessian.end(), hessian);
    if (max_asymmetry) *max_asymmetry = r.max_asymmetry;
    if (repulsion) std::copy(r.h_repulsion.begin(), r.h_repulsion.end(), repulsion);
    if (dispersion) std::copy(r.h_dispersion.begin(), r.h_dispersion.end(), dispersion);
    if (halogen) std::copy(r.h_halogen.begin(), r.h_halogen.end(), halogen);
    return GFN1_OK;
}

int gfn1_hessian_batch(int32_t n, const gfn1_system* systems, const gfn1_params* params,
                       const gfn1_electronic_options* electronic, const gfn1_hessian_options* options,
                       double* hessians, double* dipole_derivatives, int32_t* status, gfn1_error* err) {
    if ((n > 0 && (!hessians || !status))) {
        clear(err);
        return fail(err, Status::error(kErrInvalidInput, "invalid Hessian arguments"));
    }
    std::vector<HessianResult> res;
    int rc = hessian_common(n, systems, params, electronic, options, res, err);
    if (rc != GFN1_OK) return rc;
    size_t off = 0, doff = 0;
    for (int i = 0; i < n; ++i) {
        const size_t len = (size_t)9 * systems[i].nat * systems[i].nat;
        status[i] = res[i].status;
        if (res[i].status == kOk) {
            std::copy(res[i].hessian.begin(), res[i].hessian.end(), hessians + off);
            if (dipole_derivatives)
                std::copy(res[i].dipole_derivatives.begin(), res[i].dipole_derivatives.end(),
                          dipole_derivatives + doff);
        }
        off += len;
        doff += (size_t)9 * systems[i].nat;
    }
    return GFN1_OK;
}

}  // extern "C"

// --- Learned SCC start charges ---------------------------------------------

namespace {

// Systems -> concatenated atoms; instance usable when every atom is H, C, N or O
// with two shells.  nsh receives the shell count of every system.
Status to_guess_atoms(int32_t n_sys, const gfn1_system* systems, const gfn1_params* params, GuessAtoms& a,
                      std::vector<int>* nsh) {
    a = GuessAtoms{};
    a.inst_atom0.push_back(0);
    if (nsh) nsh->assign(n_sys, 0);
    for (int i = 0; i < n_sys; ++i) {
        const gfn1_system& s = systems[i];
        if (s.nat < 0 || (s.nat > 0 && (!s.numbers || !s.positions)))
            return Status::error(kErrInvalidInput, "invalid system " + std::to_string(i));
        int ok = 1;
        for (int k = 0; k < s.nat; ++k) {
            const int z = s.numbers[k];
            int ns = 0;
            if (params) {
                if (z < 1 || z > kMaxZ || !params->model.element(z).present)
                    return Status::error(kErrMissingElement, "missing element Z=" + std::to_string(z));
                ns = params->model.element(z).nshell;
            }
            if (guess_element_slot(z) < 0 || (params && ns != 2)) ok = 0;
            if (nsh) (*nsh)[i] += ns;
            a.z.push_back(z);
            a.inst.push_back(i);
            for (int d = 0; d < 3; ++d) a.x.push_back(s.positions[3 * k + d]);
        }
        a.inst_atom0.push_back((int)a.z.size());
        a.usable.push_back(ok);
    }
    return Status{};
}

// Two labels per atom from concatenated shell charges (usable systems only).
void to_labels(int32_t n_sys, const GuessAtoms& a, const std::vector<int>& nsh, const double* q,
               std::vector<float>& labels) {
    labels.assign(2 * a.z.size(), 0.0f);
    size_t off = 0;
    for (int i = 0; i < n_sys; ++i) {
        if (a.usable[i])
            for (int k = a.inst_atom0[i]; k < a.inst_atom0[i + 1]; ++k) {
                const size_t sh = off + 2 * (size_t)(k - a.inst_atom0[i]);
                labels[2 * (size_t)k] = (float)q[sh];
                labels[2 * (size_t)k + 1] = (float)q[sh + 1];
            }
        off += nsh[i];
    }
}

}  // namespace

int gfn1_params_shell_count(const gfn1_params* p, int z) {
    if (!p || z < 1 || z > kMaxZ || !p->model.element(z).present) return -1;
    return p->model.element(z).nshell;
}

int gfn1_guess_model_create(uint64_t seed, gfn1_guess_model** out, gfn1_error* err) {
    clear(err);
    if (!out) return fail(err, Status::error(kErrInvalidInput, "null output"));
    auto m = std::make_unique<gfn1_guess_model>();
    Status st = m->m.init_random(seed);
    if (!st.ok()) return fail(err, st);
    *out = m.release();
    return GFN1_OK;
}

int gfn1_guess_model_load(const char* path, gfn1_guess_model** out, gfn1_error* err) {
    clear(err);
    if (!out || !path) return fail(err, Status::error(kErrInvalidInput, "null argument"));
    auto m = std::make_unique<gfn1_guess_model>();
    Status st = m->m.load(path);
    if (!st.ok()) return fail(err, st);
    *out = m.release();
    return GFN1_OK;
}

int gfn1_guess_model_save(const gfn1_guess_model* m, const char* path, gfn1_error* err) {
    clear(err);
    if (!m || !path) return fail(err, Status::error(kErrInvalidInput, "null argument"));
    Status st = m->m.save(path);
    return st.ok() ? GFN1_OK : fail(err, st);
}

int gfn1_guess_model_get_params(const gfn1_guess_model* m, float* theta, float* norm) {
    if (!m || !theta || m->m.params().size() != (size_t)kGuessParams) return GFN1_ERR_INVALID_INPUT;
    std::memcpy(theta, m->m.params().data(), sizeof(float) * kGuessParams);
    if (norm) std::memcpy(norm, m->m.norm().data(), sizeof(float) * 2 * kGuessFeat);
    return GFN1_OK;
}

int gfn1_guess_model_set_params(gfn1_guess_model* m, const float* theta) {
    if (!m || !theta) return GFN1_ERR_INVALID_INPUT;
    return m->m.set_params(theta).ok() ? GFN1_OK : GFN1_ERR_INVALID_INPUT;
}

void gfn1_guess_model_free(gfn1_guess_model* m) { delete m; }

void gfn1_guess_train_options_default(gfn1_guess_train_options* o) {
    if (!o) return;
    GuessTrainOptions d;
    o->epochs = d.epochs;
    o->batch_atoms = d.batch_atoms;
    o->learning_rate = d.learning_rate;
    o->seed = d.seed;
}

int gfn1_guess_train(gfn1_guess_model* m, int32_t n_sys, const gfn1_system* systems, const gfn1_params* params,
                     const double* shell_charges, const gfn1_guess_train_options* options, float* loss_per_epoch,
                     gfn1_error* err) {
    clear(err);
    if (!m || n_sys <= 0 || !systems || !params || !shell_charges)
        return fail(err, Status::error(kErrInvalidInput, "invalid training arguments"));
    gfn1_guess_train_options def;
    if (!options) {
        gfn1_guess_train_options_default(&def);
        options = &def;
    }
    GuessAtoms a;
    std::vector<int> nsh;
    Status st = to_guess_atoms(n_sys, systems, params, a, &nsh);
    if (!st.ok()) return fail(err, st);
    std::vector<float> labels;
    to_labels(n_sys, a, nsh, shell_charges, labels);
    GuessTrainOptions o;
    o.epochs = options->epochs;
    o.batch_atoms = options->batch_atoms;
    o.learning_rate = options->learning_rate;
    o.seed = options->seed;
    std::vector<float> loss;
    st = m->m.train(a, labels, o, loss);
    if (!st.ok()) return fail(err, st);
    if (loss_per_epoch) std::copy(loss.begin(), loss.end(), loss_per_epoch);
    return GFN1_OK;
}

namespace {
// Start vectors from per-atom network outputs: each usable system's shell
// charges shifted to sum to its charge, zeros for the others.
void shifted_start_charges(int32_t n_sys, const gfn1_system* systems, const GuessAtoms& a, const std::vector<int>& nsh,
                           const std::vector<float>& y, double* shell_charges) {
    size_t off = 0;
    for (int i = 0; i < n_sys; ++i) {
        for (int k = 0; k < nsh[i]; ++k) shell_charges[off + k] = 0.0;
        if (a.usable[i] && nsh[i] > 0) {
            double sum = 0.0;
            for (int k = a.inst_atom0[i]; k < a.inst_atom0[i + 1]; ++k) sum += (double)y[2 * k] + (double)y[2 * k + 1];
            const double shift = (sum - systems[i].charge) / nsh[i];
            for (int k = a.inst_atom0[i]; k < a.inst_atom0[i + 1]; ++k) {
                const size_t sh = off + 2 * (size_t)(k - a.inst_atom0[i]);
                shell_charges[sh] = (double)y[2 * k] - shift;
                shell_charges[sh + 1] = (double)y[2 * k + 1] - shift;
            }
        }
        off += nsh[i];
    }
}

int guess_predict(const gfn1_guess_model* m, int32_t n_sys, const gfn1_system* systems, const gfn1_params* params,
                  double* shell_charges, gfn1_error* err, bool noalloc) {
    clear(err);
    if (!m || n_sys < 0 || (n_sys > 0 && !systems) || !params || !shell_charges)
        return fail(err, Status::error(kErrInvalidInput, "invalid prediction arguments"));
    GuessAtoms a;
    std::vector<int> nsh;
    Status st = to_guess_atoms(n_sys, systems, params, a, &nsh);
    if (!st.ok()) return fail(err, st);
    std::vector<float> y;
    st = noalloc ? m->m.predict_atoms_noalloc(a, y) : m->m.predict_atoms(a, y);
    if (!st.ok()) return fail(err, st);
    shifted_start_charges(n_sys, systems, a, nsh, y, shell_charges);
    return GFN1_OK;
}
}  // namespace

int gfn1_guess_predict(const gfn1_guess_model* m, int32_t n_sys, const gfn1_system* systems,
                       const gfn1_params* params, double* shell_charges, gfn1_error* err) {
    return guess_predict(m, n_sys, systems, params, shell_charges, err, false);
}

int gfn1_guess_predict_noalloc(const gfn1_guess_model* m, int32_t n_sys, const gfn1_system* systems,
                               const gfn1_params* params, double* shell_charges, gfn1_error* err) {
    return guess_predict(m, n_sys, systems, params, shell_charges, err, true);
}

static_assert(GFN1_NOALLOC_MAX_ATOMS == gfn1::kNoAllocMaxAtoms && GFN1_NOALLOC_MAX_SYSTEMS == gfn1::kNoAllocMaxSystems,
              "no-allocation limits of the header and the library agree");
size_t gfn1_guess_noalloc_scratch_bytes(void) { return gfn1::guess_noalloc_scratch_bytes(); }

int gfn1_guess_features(int32_t n_sys, const gfn1_system* systems, float* out, gfn1_error* err) {
    clear(err);
    if (n_sys < 0 || (n_sys > 0 && !systems) || !out)
        return fail(err, Status::error(kErrInvalidInput, "invalid arguments"));
    GuessAtoms a;
    Status st = to_guess_atoms(n_sys, systems, nullptr, a, nullptr);
    if (!st.ok()) return fail(err, st);
    std::vector<float> f;
    st = guess_features(a, f);
    if (!st.ok()) return fail(err, st);
    std::copy(f.begin(), f.end(), out);
    return GFN1_OK;
}

int gfn1_guess_loss_gradient(const gfn1_guess_model* m, int32_t n_sys, const gfn1_system* systems,
                             const gfn1_params* params, const double* shell_charges, double* loss, float* grad,
                             gfn1_error* err) {
    clear(err);
    if (!m || n_sys <= 0 || !systems || !params || !shell_charges || !loss || !grad)
        return fail(err, Status::error(kErrInvalidInput, "invalid arguments"));
    GuessAtoms a;
    std::vector<int> nsh;
    Status st = to_guess_atoms(n_sys, systems, params, a, &nsh);
    if (!st.ok()) return fail(err, st);
    std::vector<float> labels, g;
    to_labels(n_sys, a, nsh, shell_charges, labels);
    st = m->m.loss_gradient(a, labels, loss, g);
    if (!st.ok()) return fail(err, st);
    std::copy(g.begin(), g.end(), grad);
    return GFN1_OK;
}
