// C ABI of the engine (include/gfn1_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gfn1_amd.h"
#include "engine.hpp"
#include "model.hpp"

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

struct gfn1_batch {
    std::unique_ptr<Batch> batch;
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
    return Status{};
}

SystemInput to_input(const gfn1_system* s, const gfn1_electronic_options* o) {
    SystemInput in;
    in.nat = s->nat;
    in.z = s->numbers;
    in.x = s->positions;
    in.charge = (o && o->has_charge) ? o->charge : s->charge;
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
    if (options->precision != GFN1_PRECISION_FP64)
        return fail(err, Status::error(kErrInvalidInput, "only the FP64 precision mode is implemented"));
    if (options->guess != GFN1_GUESS_ZERO)
        return fail(err, Status::error(kErrInvalidInput, "only the standard (zero) charge guess is implemented"));
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
    Status st;
    const DeviceModel* dm = const_cast<gfn1_params*>(params)->device_model(st);
    if (!dm) return fail(err, st);
    auto b = std::make_unique<gfn1_batch>();
    b->batch = std::make_unique<Batch>();
    st = b->batch->create(*dm, sys, e);
    if (!st.ok()) return fail(err, st);
    *out = b.release();
    return GFN1_OK;
}

int gfn1_batch_run(gfn1_batch* batch, gfn1_error* err) {
    clear(err);
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

void gfn1_batch_timings(const gfn1_batch* b, double out[4]) {
    if (!b) return;
    std::memcpy(out, b->batch->timings(), sizeof(double) * 4);
}

void gfn1_batch_free(gfn1_batch* b) { delete b; }

}  // extern "C"
