/*
 * gfn1_amd.h -- C ABI of the MI355X GFN1-xTB SCC energy + analytic gradient engine.
 *
 * This is the drop-in boundary for the reference's non-periodic SCC path
 * (energy and analytic gradient), and for its callers built on it here: the
 * L-BFGS geometry optimiser, batched molecular dynamics and the Cartesian
 * Hessian.  Each entry point names
 * the reference interface it replaces (paths relative to the reference
 * repository root).  All coordinates are in bohr, energies in Hartree,
 * gradients in Hartree/bohr.  Matrices are dense n x n FP64; they are symmetric,
 * so row- and column-major layouts coincide.
 *
 * Ownership: results own their buffers; free them with the matching *_free.
 * Accessor pointers stay valid until the result is freed.  Thread safety: every
 * call is re-entrant; concurrent calls use separate device streams.  A batch,
 * optimisation or result handle is used by one thread at a time.
 */
#ifndef GFN1_AMD_H
#define GFN1_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes: one per Gfn1Error variant (src/error.rs:8-17), plus device errors. */
enum gfn1_status {
    GFN1_OK = 0,
    GFN1_ERR_IO = 1,
    GFN1_ERR_PARSE = 2,
    GFN1_ERR_INVALID_INPUT = 3,
    GFN1_ERR_MISSING_ELEMENT = 4,
    GFN1_ERR_MISSING_GLOBAL = 5,
    GFN1_ERR_LINEAR_ALGEBRA = 6,
    GFN1_ERR_SINGULAR_CELL = 7,
    GFN1_ERR_SCC_NOT_CONVERGED = 8,
    GFN1_ERR_DEVICE = 100
};

typedef struct gfn1_error {
    int code;
    int iterations;   /* SccNotConverged{iterations, rms} payload */
    double rms;
    char message[256];
} gfn1_error;

typedef struct gfn1_params gfn1_params;

/* Gfn1Parameters::resolve (src/params.rs:204-230).  spec: NULL or "" ->
 * $GFN1_XTB_PARAM, else the bundled GFN1-xTB file; "builtin" / "builtin:gfn1"
 * -> the bundled GFN1-xTB file; "builtin:si" / "builtin:gfn1-si" / "builtin-si"
 * -> the bundled GFN1(Si)-xTB file; anything else is a parameter-file path. */
int gfn1_params_load(const char* spec, gfn1_params** out, gfn1_error* err);
/* Gfn1Parameters::from_str (src/params.rs:272): parameters from the text of a
 * parameter file (source "inline text"). */
int gfn1_params_from_text(const char* text, gfn1_params** out, gfn1_error* err);
void gfn1_params_free(gfn1_params* params);
/* Gfn1Parameters::source_description (src/params.rs:236-270): writes at most
 * len - 1 characters and a terminating NUL into out (nullable when len is 0);
 * returns the full length of the description. */
int32_t gfn1_params_source(const gfn1_params* params, char* out, int32_t len);
/* Number of shells of element z in the loaded parameters (-1 if absent).  The
 * shell layout of a system is its atoms' shells in input order
 * (src/basis.rs:121-233). */
int gfn1_params_shell_count(const gfn1_params* params, int z);

/* PeriodicSystem (src/system.rs:17-21).  lattice: NULL for a molecule, or the
 * three cell vectors a, b, c of a fully periodic cell in bohr, one after the
 * other (9 doubles).  A system with a lattice takes the periodic Gamma-point
 * path, as run_electronic and analytic_gradient dispatch it
 * (src/electronic.rs:829-832, src/gradient.rs:61-66): image sums of S and H0
 * within 30 bohr, the Ewald-summed shell Coulomb matrix, lattice-summed CN,
 * repulsion, D3 and halogen terms, no level shift and no fallback ladder in
 * the SCC, and every gradient term whatever the include flags.  A cell with
 * |det| < 1e-14 fails with GFN1_ERR_SINGULAR_CELL. */
typedef struct gfn1_system {
    int32_t nat;
    const int32_t* numbers;   /* atomic numbers, nat */
    const double* positions;  /* bohr, 3*nat, atom-major xyz */
    double charge;
    const double* lattice;    /* NULL, or 9 doubles: a, b, c (bohr) */
} gfn1_system;

/* PeriodicSystem::from_xyz_str (src/system.rs:45-99): line 1 the atom count,
 * line 2 a comment, then "symbol x y z" per atom (a symbol in any letter case,
 * or an atomic number 1..86).  Coordinates are Angstrom unless
 * coordinates_are_bohr.  An extended-XYZ comment Lattice="ax ay az bx by bz cx
 * cy cz" gives the system a lattice (same units as the coordinates;
 * src/system.rs:191-219); a pbc="T T T" field, when present, must mark all
 * three directions periodic (only fully periodic cells are supported).  Errors:
 * GFN1_ERR_INVALID_INPUT for the count, a short or missing line, an unknown
 * element or a malformed Lattice/pbc field; GFN1_ERR_PARSE for a coordinate
 * that is not a number (the message names the line); GFN1_ERR_SINGULAR_CELL
 * for cell vectors with |det| < 1e-14. */
typedef struct gfn1_xyz gfn1_xyz;
int gfn1_system_from_xyz(const char* text, double charge, int32_t coordinates_are_bohr, gfn1_xyz** out,
                         gfn1_error* err);
/* The parsed system (positions in bohr); valid until gfn1_xyz_free. */
const gfn1_system* gfn1_xyz_system(const gfn1_xyz* xyz);
void gfn1_xyz_free(gfn1_xyz* xyz);

/* In-scope subset of ElectronicOptions (src/electronic.rs:157-393) with the same
 * defaults (:395-452), plus HamiltonianOptions (src/hamiltonian.rs:16-33). */
typedef struct gfn1_electronic_options {
    int32_t has_charge;           /* Option<f64> charge: 0 -> use system charge */
    double charge;
    int32_t max_scc;              /* 250 */
    double energy_tolerance;      /* 1e-6 */
    double charge_tolerance;      /* 2e-5 */
    double mixing;                /* 0.4 */
    int32_t scc_broyden;          /* 1 */
    int32_t scc_broyden_size;     /* 250 */
    double level_shift;           /* 0 */
    int32_t nprim;                /* 0 = GFN1 primitive-count rule; only 0 is supported */
    double eigen_tolerance;       /* 1e-12 */
    double electronic_temperature;/* 300 K */
    int32_t enable_dispersion;    /* 1 */
    double integral_cutoff;       /* 30 bohr */
    double coordination_cutoff;   /* 30 bohr */
    int32_t enable_cn_hamiltonian;/* 1 */
    const double* scc_initial_shell_charges; /* nullable; used when its length is nsh */
    int32_t scc_initial_shell_charges_len;
} gfn1_electronic_options;

void gfn1_electronic_options_default(gfn1_electronic_options* options);

/* AnalyticGradientOptions (src/gradient.rs:21-41). */
typedef struct gfn1_gradient_options {
    gfn1_electronic_options electronic;
    int32_t include_repulsion;
    int32_t include_dispersion;
    int32_t include_hamiltonian;
    int32_t include_scc;
    int32_t include_halogen;
} gfn1_gradient_options;

void gfn1_gradient_options_default(gfn1_gradient_options* options);

typedef struct gfn1_electronic_result gfn1_electronic_result;
typedef struct gfn1_gradient_result gfn1_gradient_result;

/* run_electronic (src/electronic.rs:800-847). */
int gfn1_run_electronic(const gfn1_system* system, const gfn1_params* params,
                        const gfn1_electronic_options* options,
                        gfn1_electronic_result** out, gfn1_error* err);

/* analytic_gradient (src/gradient.rs:56-69): run_electronic followed by
 * analytic_gradient_from_result (src/gradient.rs:122-242). */
int gfn1_analytic_gradient(const gfn1_system* system, const gfn1_params* params,
                           const gfn1_gradient_options* options,
                           gfn1_gradient_result** out, gfn1_error* err);

void gfn1_electronic_result_free(gfn1_electronic_result* result);
void gfn1_gradient_result_free(gfn1_gradient_result* result);

/* --- ElectronicResult accessors (src/electronic.rs:600-654) --- */
void gfn1_electronic_result_sizes(const gfn1_electronic_result* r, int32_t* nat, int32_t* nsh, int32_t* nao);

/* EnergyTerms::named_values order (src/electronic.rs:676-694):
 * total_free, total_internal, repulsion, electronic, isotropic_scc, third_order,
 * higher_order, dispersion, halogen, multipole, exchange, spin_polarization,
 * plus_u, external_field, electronic_entropy. */
void gfn1_electronic_result_energy_terms(const gfn1_electronic_result* r, double out[15]);

enum gfn1_matrix_field {
    GFN1_MAT_OVERLAP = 0,
    GFN1_MAT_H0 = 1,
    GFN1_MAT_FOCK = 2,
    GFN1_MAT_DENSITY = 3,
    GFN1_MAT_ENERGY_WEIGHTED_DENSITY = 4
};
/* nao x nao; NULL for an unknown field. */
const double* gfn1_electronic_result_matrix(const gfn1_electronic_result* r, int field);

enum gfn1_vector_field {
    GFN1_VEC_ORBITAL_ENERGIES = 0,   /* nao */
    GFN1_VEC_OCCUPATIONS = 1,        /* nao */
    GFN1_VEC_SHELL_CHARGES = 2,      /* nsh */
    GFN1_VEC_ATOMIC_CHARGES = 3,     /* nat */
    GFN1_VEC_SHELL_SCC_POTENTIAL = 4,/* nsh */
    GFN1_VEC_COORDINATION_NUMBERS = 5, /* nat */
    GFN1_VEC_DIPOLE = 6              /* 3 */
};
const double* gfn1_electronic_result_vector(const gfn1_electronic_result* r, int field);

enum gfn1_scalar_field {
    GFN1_SCALAR_ELECTRONIC_TEMPERATURE = 0,
    GFN1_SCALAR_FERMI_LEVEL = 1,
    GFN1_SCALAR_NELEC = 2,
    GFN1_SCALAR_ITERATIONS = 3,
    GFN1_SCALAR_CONVERGED = 4,
    GFN1_SCALAR_CHARGE_ORDER = 5
};
double gfn1_electronic_result_scalar(const gfn1_electronic_result* r, int field);

/* --- AnalyticGradientResult accessors (src/gradient.rs:44-54) --- */
const gfn1_electronic_result* gfn1_gradient_result_electronic(const gfn1_gradient_result* r);
double gfn1_gradient_result_total_energy(const gfn1_gradient_result* r);
double gfn1_gradient_result_max_gradient(const gfn1_gradient_result* r);
enum gfn1_gradient_field {
    GFN1_GRAD_TOTAL = 0,
    GFN1_GRAD_FORCES = 1,
    GFN1_GRAD_ELECTRONIC = 2,
    GFN1_GRAD_REPULSION = 3,
    GFN1_GRAD_DISPERSION = 4,
    GFN1_GRAD_HALOGEN = 5
};
/* nat x 3 (atom-major xyz). */
const double* gfn1_gradient_result_field(const gfn1_gradient_result* r, int field);

/* --- Batched energy + gradient of many independent systems ---
 * No reference counterpart as a single call: the reference runs analytic_gradient
 * once per system (src/gradient.rs:56).  Each instance follows exactly that
 * path; instances run together on the device. */
enum gfn1_precision { GFN1_PRECISION_FP64 = 0, GFN1_PRECISION_MIXED = 1 };
enum gfn1_guess { GFN1_GUESS_ZERO = 0, GFN1_GUESS_LEARNED = 1 };

typedef struct gfn1_batch_options {
    gfn1_gradient_options gradient;
    int32_t precision;       /* gfn1_precision */
    int32_t guess;           /* gfn1_guess */
    int32_t eigensolver;     /* 0 auto (own kernels up to 512 AO, rocSOLVER syevd above),
                                1 rocSOLVER syevd, 2 rocSOLVER syevj, 3 rocSOLVER syevdj */
    int32_t keep_matrices;   /* copy S/H0/F/P/W back to the host per instance */
    int32_t scc_graphs;      /* 1 (default): buckets of the own eigensolver iterate by
                                replays of a captured HIP graph of one SCC iteration;
                                0: one host-launched kernel sequence per iteration.
                                Results are the same either way. */
} gfn1_batch_options;

void gfn1_batch_options_default(gfn1_batch_options* options);

typedef struct gfn1_batch gfn1_batch;

/* Upload the systems and build the per-instance index tables on the device. */
int gfn1_batch_create(int32_t n_inst, const gfn1_system* systems, const gfn1_params* params,
                      const gfn1_batch_options* options, gfn1_batch** out, gfn1_error* err);
/* Run the whole energy + gradient pass on device-resident inputs.  Blocks until
 * the results are complete on the device. */
int gfn1_batch_run(gfn1_batch* batch, gfn1_error* err);
/* Per-instance results after gfn1_batch_run. */
int gfn1_batch_status(const gfn1_batch* batch, int32_t inst);           /* gfn1_status of the instance */
double gfn1_batch_energy(const gfn1_batch* batch, int32_t inst);        /* total_free */
void gfn1_batch_energy_terms(const gfn1_batch* batch, int32_t inst, double out[15]);
int32_t gfn1_batch_iterations(const gfn1_batch* batch, int32_t inst);
/* Copies nat*3 doubles into out. */
int gfn1_batch_gradient(const gfn1_batch* batch, int32_t inst, double* out);
/* Copies nsh doubles (converged shell charges) into out. */
int gfn1_batch_shell_charges(const gfn1_batch* batch, int32_t inst, double* out);
/* Device time of the last gfn1_batch_run phases, milliseconds: [setup, scc, gradient, total]. */
void gfn1_batch_timings(const gfn1_batch* batch, double out[4]);

/* Device-side timing of one kernel over the last gfn1_batch_run: summed launch
 * durations (HIP events recorded on the stream each launch runs on), number of
 * launches, and algorithmic bytes of those launches.  For
 * GFN1_KERNEL_TRIDIAG the algorithmic bytes of an n x n matrix are one pass
 * over the trailing lower triangle per column step, 8 (n-1) n (n+1) / 6.
 * No reference counterpart (measurement hook of the batch engine). */
enum { GFN1_KERNEL_TRIDIAG = 0 };
int gfn1_batch_kernel_stats(const gfn1_batch* batch, int kernel, double* total_ms, int64_t* launches,
                            double* bytes);
/* How the last gfn1_batch_run drove the device in its SCC phase: kernels run
 * by the SCC iterations (a graph replay counts its nodes), host calls that
 * enqueued them (a kernel launch, a rocBLAS call, a rocSOLVER eigensolve or a
 * graph replay each count once), graph replays, host-launched iterations, reads
 * of a bucket's gathered states by the host (one per finished iteration or
 * replay), graphs captured (two per bucket, kept for later runs) and the host
 * time spent capturing them (microseconds).  No reference counterpart
 * (measurement hook of the batch engine). */
enum {
    GFN1_LAUNCH_KERNELS = 0,
    GFN1_LAUNCH_ENQUEUES,
    GFN1_LAUNCH_GRAPH_REPLAYS,
    GFN1_LAUNCH_HOST_ITERATIONS,
    GFN1_LAUNCH_STATE_READS,
    GFN1_LAUNCH_GRAPHS_CAPTURED,
    GFN1_LAUNCH_CAPTURE_US,
    GFN1_LAUNCH_STATS
};
int gfn1_batch_launch_stats(const gfn1_batch* batch, int64_t out[GFN1_LAUNCH_STATS]);
/* Autotuning of the eigenvector back-transformation: after a gfn1_batch_run,
 * times each admissible width (16, 32, 64 eigenvectors per work-group) on every
 * size bucket's reflectors from that run and keeps the fastest for the later
 * runs.  All widths give bit-identical results.  gfn1_batch_tuning_entry
 * returns, for tuned bucket k < gfn1_batch_tuning_count: out_i = {padded
 * dimension, instances timed, chosen width}, out_ms = mean launch time per
 * width 16, 32, 64 (-1 where not admissible).  No reference counterpart. */
int gfn1_batch_autotune(gfn1_batch* batch, gfn1_error* err);
/* Fixes the back-transformation width (16, 32 or 64; 64 only for buckets of
 * at most 256 rows, the others keep theirs) in place of the tuned one. */
int gfn1_batch_set_backtransform_width(gfn1_batch* batch, int32_t cols, gfn1_error* err);
int32_t gfn1_batch_tuning_count(const gfn1_batch* batch);
int gfn1_batch_tuning_entry(const gfn1_batch* batch, int32_t k, int32_t out_i[3], double out_ms[3]);
void gfn1_batch_free(gfn1_batch* batch);

/* --- Learned SCC start charges ---
 * A per-atom MLP (68 radial-environment inputs, two hidden layers of 64 SiLU
 * units, two outputs: the atom's two shell charges) predicts the SCC start
 * vector; per system the predictions are shifted to sum to the system charge.
 * The start vector enters through the reference's scc_initial_shell_charges
 * hook (src/electronic.rs:386-392, :1231-1239); the converged result does not
 * depend on it.  Systems with atoms other than H, C, N, O keep the standard
 * start.  No reference counterpart otherwise.  Training and inference run on
 * the GPU with the library's own kernels (FP32, Adam). */
enum { GFN1_GUESS_PARAMS = 8706, GFN1_GUESS_FEATURES = 68 };
typedef struct gfn1_guess_model gfn1_guess_model;
int gfn1_guess_model_create(uint64_t seed, gfn1_guess_model** out, gfn1_error* err); /* random initialisation */
int gfn1_guess_model_load(const char* path, gfn1_guess_model** out, gfn1_error* err);
int gfn1_guess_model_save(const gfn1_guess_model* model, const char* path, gfn1_error* err);
/* theta: GFN1_GUESS_PARAMS floats; norm (nullable): descriptor mean[68] then inverse spread[68]. */
int gfn1_guess_model_get_params(const gfn1_guess_model* model, float* theta, float* norm);
int gfn1_guess_model_set_params(gfn1_guess_model* model, const float* theta);
void gfn1_guess_model_free(gfn1_guess_model* model);

typedef struct gfn1_guess_train_options {
    int32_t epochs;        /* passes over the training atoms (default 40) */
    int32_t batch_atoms;   /* atoms per Adam step (default 1024) */
    float learning_rate;   /* Adam step size (default 1e-3) */
    uint64_t seed;         /* shuffling */
} gfn1_guess_train_options;
void gfn1_guess_train_options_default(gfn1_guess_train_options* options);
/* shell_charges: converged shell charges of every system, concatenated (nsh per
 * system).  loss_per_epoch (nullable): epochs mean squared errors. */
int gfn1_guess_train(gfn1_guess_model* model, int32_t n_sys, const gfn1_system* systems, const gfn1_params* params,
                     const double* shell_charges, const gfn1_guess_train_options* options, float* loss_per_epoch,
                     gfn1_error* err);
/* Start vectors (nsh per system, concatenated). */
int gfn1_guess_predict(const gfn1_guess_model* model, int32_t n_sys, const gfn1_system* systems,
                       const gfn1_params* params, double* shell_charges, gfn1_error* err);
/* gfn1_guess_predict without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and results
 * as gfn1_guess_predict; the library's device memory for the call is a static
 * array of gfn1_guess_noalloc_scratch_bytes() bytes held by every device the
 * library runs on until the process ends.  Limits: at most
 * GFN1_NOALLOC_MAX_ATOMS atoms in GFN1_NOALLOC_MAX_SYSTEMS systems
 * (GFN1_ERR_INVALID_INPUT beyond them), and the model's parameters must be on
 * the device already (they are once the model is created or loaded with a
 * device present).  Calls in one process share the array and are serialised;
 * inputs and outputs are host arrays. */
enum { GFN1_NOALLOC_MAX_ATOMS = 16384, GFN1_NOALLOC_MAX_SYSTEMS = 16384 };
int gfn1_guess_predict_noalloc(const gfn1_guess_model* model, int32_t n_sys, const gfn1_system* systems,
                               const gfn1_params* params, double* shell_charges, gfn1_error* err);
size_t gfn1_guess_noalloc_scratch_bytes(void);
/* Descriptors of every atom (68 floats per atom, systems concatenated). */
int gfn1_guess_features(int32_t n_sys, const gfn1_system* systems, float* out, gfn1_error* err);
/* Training loss (mean squared shell-charge error over covered atoms) and its
 * gradient with respect to the parameters, full batch. */
int gfn1_guess_loss_gradient(const gfn1_guess_model* model, int32_t n_sys, const gfn1_system* systems,
                             const gfn1_params* params, const double* shell_charges, double* loss, float* grad,
                             gfn1_error* err);
/* Use the model for the start vectors of the batch's next runs (nullptr: back
 * to the standard start).  A batch created with GFN1_GUESS_LEARNED needs a model
 * before gfn1_batch_run. */
int gfn1_batch_set_guess(gfn1_batch* batch, const gfn1_guess_model* model, gfn1_error* err);

/* New positions (bohr, atom-major xyz; the atoms of every system in batch
 * order) for the batch's next runs.  All coordinates must be finite.  No
 * reference counterpart as a call: the reference builds a new system per
 * geometry (src/optimizer.rs:316-322). */
int gfn1_batch_set_positions(gfn1_batch* batch, const double* positions, gfn1_error* err);

/* All instances at once, systems in batch order: statuses (n), total free
 * energies (n; NaN where the status is not GFN1_OK), gradients (3 nat each) and
 * converged shell charges (nsh each), after gfn1_batch_run. */
int gfn1_batch_statuses(const gfn1_batch* batch, int32_t* out);
int gfn1_batch_energies(const gfn1_batch* batch, double* out);
int gfn1_batch_gradients(const gfn1_batch* batch, double* out);
int gfn1_batch_all_shell_charges(const gfn1_batch* batch, double* out);
/* Start shell charges of every instance (nsh each, batch order) for the next runs. */
int gfn1_batch_set_all_start_charges(gfn1_batch* batch, const double* shell_charges, gfn1_error* err);

/* Start shell charges (nsh values) of one instance for the batch's next runs,
 * through the reference's scc_initial_shell_charges hook
 * (src/electronic.rs:386-392, :1231-1239); for drivers that start each SCC
 * from the previous geometry's charges.  No reference counterpart as a call. */
int gfn1_batch_set_start_charges(gfn1_batch* batch, int32_t inst, const double* shell_charges, gfn1_error* err);

/* Current positions of every atom (bohr, 3 per atom, batch order). */
int gfn1_batch_positions(const gfn1_batch* batch, double* out);

/* --- Molecular dynamics ---
 * NVE velocity Verlet for every system of the batch at once.  The reference
 * runs dynamics through ASE with its calculator (no library call); each step
 * here is one batched energy + gradient pass over the systems still running. */
typedef struct gfn1_md_options {
    double timestep_fs;   /* 0.5 */
    int32_t steps;        /* 10 */
    int32_t warm_start;   /* 1: each SCC starts from the previous step's charges */
} gfn1_md_options;
void gfn1_md_options_default(gfn1_md_options* options);
/* From the batch's current positions.  velocities: 3 per atom (bohr per atomic
 * time unit), updated in place; potential / kinetic (nullable): (steps + 1) x n
 * energies (Eh), row = step, NaN once a system has stopped; status (nullable):
 * n gfn1_status values; scc_iterations (nullable): (steps + 1) x n SCC
 * iteration counts, -1 once a system has stopped.  With warm_start the
 * batch's learned start-charge model (if any) starts step 0 only; the batch's
 * model and start charges are restored when the call returns.  Afterwards the
 * batch holds the final positions (gfn1_batch_positions) and no batch results. */
int gfn1_batch_md(gfn1_batch* batch, const gfn1_md_options* options, double* velocities, double* potential,
                  double* kinetic, int32_t* status, int32_t* scc_iterations, gfn1_error* err);

/* --- Geometry optimisation ---
 * optimize_geometry (src/optimizer.rs:78-185): L-BFGS on the total free energy
 * and the analytic gradient, Armijo backtracking line search, per-atom step
 * cap.  Non-periodic systems only. */
typedef struct gfn1_optimization_options {
    int32_t max_iterations;      /* 250 */
    double gradient_tolerance;   /* 1e-4, on the largest gradient component */
    double step_tolerance;       /* 1e-7 bohr */
    int32_t history;             /* 12 correction pairs */
    double initial_step;         /* 1.0, first line-search step length */
    double max_atom_step;        /* 0.30 bohr; <= 0 disables the cap */
    int32_t warm_start;          /* 0: every evaluation starts its SCC as the
                                    batch is configured (the reference's
                                    behaviour); 1: from the converged shell
                                    charges of the last accepted geometry */
    int32_t keep_trajectory;     /* 1: record every accepted step */
} gfn1_optimization_options;
/* GeometryOptimizationOptions::default (src/optimizer.rs:32-44). */
void gfn1_optimization_options_default(gfn1_optimization_options* options);

typedef struct gfn1_optimization gfn1_optimization;

/* optimize_geometry for every system of the batch at once, each from its
 * current positions and with its own L-BFGS state; every round evaluates the
 * pending geometries of all systems (start points or line-search trials) in
 * one device pass over just those systems.  The batch's options (precision,
 * charge guess, gradient terms) apply to every evaluation.  A system that
 * fails (its start geometry does not evaluate, or a line search finds no
 * downhill step in 24 halvings) gets the error optimize_geometry would have
 * returned as its status; the others are unaffected.  Afterwards the batch
 * holds the last evaluated geometries and no batch results (run it again
 * for those). */
int gfn1_batch_optimize(gfn1_batch* batch, const gfn1_optimization_options* options, gfn1_optimization** out,
                        gfn1_error* err);
/* optimize_geometry of one system (src/optimizer.rs:78-82): returns the
 * system's error code when it fails, as the reference returns Err. */
int gfn1_optimize_geometry(const gfn1_system* system, const gfn1_params* params,
                           const gfn1_gradient_options* gradient_options, const gfn1_optimization_options* options,
                           gfn1_optimization** out, gfn1_error* err);
/* GeometryOptimizationResult fields (src/optimizer.rs:57-67) per system. */
int32_t gfn1_optimization_count(const gfn1_optimization* r);
int gfn1_optimization_status(const gfn1_optimization* r, int32_t inst);       /* gfn1_status */
const char* gfn1_optimization_message(const gfn1_optimization* r, int32_t inst);
double gfn1_optimization_energy(const gfn1_optimization* r, int32_t inst);     /* total_free */
double gfn1_optimization_max_gradient(const gfn1_optimization* r, int32_t inst);
int32_t gfn1_optimization_iterations(const gfn1_optimization* r, int32_t inst);
int32_t gfn1_optimization_converged(const gfn1_optimization* r, int32_t inst);
/* Copy nat*3 doubles: final positions (bohr) / gradient (Hartree/bohr). */
int gfn1_optimization_positions(const gfn1_optimization* r, int32_t inst, double* out);
int gfn1_optimization_gradient(const gfn1_optimization* r, int32_t inst, double* out);
/* Trajectory (GeometryOptimizationStep, src/optimizer.rs:47-55): step 0 is the
 * start geometry; positions (nullable) receives nat*3 doubles. */
int32_t gfn1_optimization_trajectory_length(const gfn1_optimization* r, int32_t inst);
int gfn1_optimization_trajectory_step(const gfn1_optimization* r, int32_t inst, int32_t step, int32_t* iteration,
                                      double* energy, double* max_gradient, double* step_norm, double* positions);
/* Batch runs, energy + gradient evaluations over all systems, and wall time. */
void gfn1_optimization_stats(const gfn1_optimization* r, int32_t* rounds, int64_t* evaluations, double* total_ms);
void gfn1_optimization_free(gfn1_optimization* r);

/* --- Cartesian Hessian ---
 * The reference computes analytic_hessian(...).hessian (src/hessian/mod.rs:112-134,
 * AnalyticHessianResult :55-65) and validates it against central differences of
 * the analytic gradient at tight SCC tolerances, within 1e-6 Eh/bohr^2
 * (tests/hessian.rs:691-760).  This is that construction, built on the device:
 * the 6 nat displaced geometries of every system run as instances of the batch
 * engine (FP64).  Of the analytic result's component breakdown, the
 * repulsion, dispersion and halogen parts come from the same differences of
 * those terms' own gradients (gfn1_hessian_terms); the electronic parts
 * (fixed_scc, fixed_pulay, fixed_cn_h0, cpxtb_response) are one remainder.  The
 * include flags select the differentiated gradient terms. */
typedef struct gfn1_hessian_options {
    double step;                  /* 1e-4 bohr */
    double scc_energy_tolerance;  /* 1e-11, replaces the electronic options' value */
    double scc_charge_tolerance;  /* 1e-11, likewise */
    int32_t include_repulsion;    /* 1 */
    int32_t include_electronic;   /* 1: the reference's include_fixed_scc, include_fixed_pulay,
                                     include_fixed_cn_h0 and include_electronic together */
    int32_t include_dispersion;   /* 1 */
    int32_t include_halogen;      /* 1 */
    int32_t warm_start;           /* 1: displaced SCCs start from the undisplaced geometry's charges
                                     (5-6 SCC iterations instead of 9-10; needs the 1e-11 charge tolerance) */
    int32_t max_batch;            /* 4096 displaced geometries per device pass */
} gfn1_hessian_options;
void gfn1_hessian_options_default(gfn1_hessian_options* options);
/* One system: hessian receives (3 nat)^2 doubles, row-major, index 3*atom + axis,
 * symmetrised; max_asymmetry (nullable) the largest |H_ij - H_ji| before that.
 * The same displaced evaluations give the Cartesian dipole derivatives
 * (dipole_derivatives, nullable: 3 nat x 3 doubles, d mu_alpha / d R_c at
 * [3 c + alpha], e) and the Mulliken dipole (dipole, nullable: 3 doubles, e bohr,
 * origin 0), the quantities of dipole_derivatives (src/properties.rs:80-145). */
int gfn1_hessian(const gfn1_system* system, const gfn1_params* params, const gfn1_electronic_options* electronic,
                 const gfn1_hessian_options* options, double* hessian, double* max_asymmetry,
                 double* dipole_derivatives, double* dipole, gfn1_error* err);
/* One system, with the classical parts of the breakdown: repulsion, dispersion
 * and halogen (each nullable, (3 nat)^2 doubles in the layout of hessian,
 * symmetrised) are the repulsion / dispersion / halogen sub-results of
 * AnalyticHessianResult (src/hessian/mod.rs:55-65: RepulsionHessianResult,
 * DispersionHessianResult, HalogenHessianResult .hessian), by central
 * differences of the terms' analytic gradients; zero for a term whose include
 * flag is off.  hessian minus the three is the electronic part; hessian and
 * max_asymmetry (nullable) as in gfn1_hessian. */
int gfn1_hessian_terms(const gfn1_system* system, const gfn1_params* params, const gfn1_electronic_options* electronic,
                       const gfn1_hessian_options* options, double* hessian, double* max_asymmetry,
                       double* repulsion, double* dispersion, double* halogen, gfn1_error* err);
/* Many systems: hessians receives the (3 nat_i)^2 blocks one after another,
 * dipole_derivatives (nullable) the 9 nat_i blocks, status (n entries) each
 * system's gfn1_status (its blocks are left untouched unless GFN1_OK). */
int gfn1_hessian_batch(int32_t n, const gfn1_system* systems, const gfn1_params* params,
                       const gfn1_electronic_options* electronic, const gfn1_hessian_options* options,
                       double* hessians, double* dipole_derivatives, int32_t* status, gfn1_error* err);

/* Library version string. */
const char* gfn1_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GFN1_AMD_H */
