// Batched GFN1-xTB energy + gradient driver (host side of the device engine).
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <string>
#include <vector>

#include "device.hpp"
#include "eigen.hpp"
#include "guess.hpp"
#include "tridiag_large.hpp"
#include "kernels.hpp"
#include "model.hpp"
#include "periodic.hpp"

namespace gfn1 {

// Model tables resident on one device.
class DeviceModel {
   public:
    Status init(const Model& m);
    ~DeviceModel();
    DeviceTables tables() const;
    const Model& model() const { return model_; }

   private:
    Model model_;
    ShellType* shells_ = nullptr;
    ElementType* elements_ = nullptr;
    double* c6_ = nullptr;
    double* pair_ = nullptr;
};

struct SystemInput {
    int nat = 0;
    const int32_t* z = nullptr;
    const double* x = nullptr;  // bohr
    double charge = 0.0;
    const double* q0 = nullptr;  // optional initial shell charges
    int q0_len = 0;
    // periodic cell: three cell vectors (bohr), one after the other; nullptr
    // for a molecule.  A periodic instance takes the Gamma-point path
    // (pbc/mod.rs:83-91).
    const double* cell = nullptr;
};

// Host-side work of one run's SCC phase.  kernels counts the kernels the SCC
// iterations ran (a graph replay contributes its node count); enqueues counts
// the host calls that put them on a stream (one per kernel launch, rocBLAS call
// or rocSOLVER eigensolve, one per graph replay); state_reads counts the times the host read a
// bucket's gathered states, one per finished iteration or replay.
struct LaunchStats {
    long long kernels = 0, enqueues = 0, graph_replays = 0, host_iterations = 0, state_reads = 0;
    long long graphs_captured = 0;
    double capture_ms = 0.0;  // host time spent capturing and instantiating graphs
};

struct EngineOptions {
    // electronic
    int max_scc = 250;
    double energy_tolerance = 1e-6;
    double charge_tolerance = 2e-5;
    double mixing = 0.4;
    int scc_broyden = 1;
    int scc_broyden_size = 250;
    double level_shift = 0.0;
    double eigen_tolerance = 1e-12;
    double electronic_temperature = 300.0;
    int enable_dispersion = 1;
    double integral_cutoff = 30.0;
    double coordination_cutoff = 30.0;
    int enable_cn_hamiltonian = 1;
    // gradient
    int compute_gradient = 1;
    int include_repulsion = 1, include_dispersion = 1, include_hamiltonian = 1, include_scc = 1,
        include_halogen = 1;
    // execution
    int eigensolver = 0;   // 0 auto (own up to 512 AO, syevd above), 1 syevd, 2 syevj, 3 syevdj
    int keep_matrices = 0;
    // iterate buckets of the own eigensolver by replays of a captured iteration
    // (0: one host-launched kernel sequence per iteration)
    int scc_graphs = 1;
    // mixed precision: FP32 MFMA for the energy-weighted density W and FP32
    // gradient contraction with compensated per-atom sums; the eigensolve, the
    // density P, Broyden mixing and energy sums stay FP64
    int mixed = 0;
};

// Per-instance results on the host.
struct InstanceResult {
    int status = 0;
    std::string message;
    int iterations = 0;
    int converged = 0;
    double energies[15] = {};
    double fermi_level = 0, temperature = 0, nelec = 0, dipole[3] = {};
    double rms = 0;
    std::vector<double> gradient, g_el, g_rep, g_disp, g_hal;
    std::vector<double> shell_charges, atomic_charges, shell_potential, cn, eps, occ;
    std::vector<double> S, H0, F, P, W;  // n x n, only with keep_matrices
    int nat = 0, nsh = 0, nao = 0;
};

class Batch {
   public:
    Batch() = default;
    ~Batch();
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    Status create(const DeviceModel& dm, const std::vector<SystemInput>& systems, const EngineOptions& opt);
    // select: nullptr runs every instance; otherwise only instances with a
    // nonzero entry are set up, iterated and differentiated (the others keep
    // status kSkipped in this run's results).
    Status run(const unsigned char* select = nullptr);
    Status download(std::vector<InstanceResult>& out);
    // New positions (bohr, atom-major xyz of every atom in instance order) for
    // the next runs; every coordinate must be finite.
    Status set_positions(const double* x);
    // Start shell charges of one instance for the next runs (nsh values).
    Status set_start_charges(int inst, const double* q);
    // Total free energy, gradient (3 * natoms, instance order), shell charges
    // (nullable; nshells of every instance, instance order) and per-instance
    // status of the last run, without the full result download.
    // term_gradients (nullable): the repulsion, D3 and halogen parts of the
    // gradient, 3 * natoms each, one block after the other (zero where a term
    // is not part of the gradient).
    Status fetch_energy_gradient(std::vector<int>& status, std::vector<double>& energy, std::vector<double>& gradient,
                                 std::vector<double>* shell_charges, std::vector<double>* term_gradients = nullptr);
    // Current positions of every atom (bohr, 3 * natoms()).
    Status positions(double* x) const {
        std::copy(h_x_.begin(), h_x_.end(), x);
        return Status{};
    }
    // Mulliken dipole of every instance (e bohr, origin 0, 3 per instance) from
    // the charges of the last run, as the full download computes it.
    Status fetch_dipoles(std::vector<double>& dipoles);
    int natoms() const { return natoms_; }
    int atom_number(int atom) const { return h_atom_z_[atom]; }  // Z of an atom (batch-wide index)
    int atom_offset(int inst) const { return h_atom0_[inst]; }   // inst in [0, size()]
    int shell_count(int inst) const { return h_nsh_[inst]; }
    int shell_offset(int inst) const { return h_sh0_[inst]; }    // into fetch_energy_gradient's shell charges
    int padded_dimension(int inst) const { return h_ld_[inst]; }  // ld of the instance's matrices
    const std::string& message(int inst) const { return h_errmsg_[inst]; }
    const double* timings() const { return timings_; }
    int size() const { return ninst_; }
    // SCC iterations of an instance in the last run (valid after fetch_energy_gradient or download)
    int iterations(int inst) const { return h_state_.empty() ? 0 : h_state_[inst].iter; }
    // Learned start charges for the next runs (nullptr: the caller's or zero start).
    Status set_guess(const GuessModel* model);
    const GuessModel* guess() const { return guess_; }
    // Start charges of every shell (padded layout), to put them back later.
    std::vector<double> start_charges() const { return h_q0_; }
    void restore_start_charges(const std::vector<double>& q0) {
        if (q0.size() == h_q0_.size() && q0 != h_q0_) {
            h_q0_ = q0;
            q0_dirty_ = true;
        }
    }
    // Device time of the tridiagonal reduction over the last run(): summed launch
    // durations (HIP events on each bucket's stream), launches, and algorithmic
    // bytes (one pass over the trailing lower triangle per column step).
    void tridiag_stats(double* ms, long long* launches, double* bytes) const {
        *ms = tri_ms_;
        *launches = tri_launches_;
        *bytes = tri_bytes_;
    }

    // How the last run's SCC phase drove the device.
    const LaunchStats& launch_stats() const { return launch_stats_; }
    // Times every back-transformation width on each own-eigensolver bucket's
    // reflectors from the last run and keeps the fastest for the later runs.
    // The widths give identical results, so tuning changes speed only.
    Status autotune();
    struct TuneRecord {
        int ld = 0, count = 0, cols = 0;
        double ms[3] = {-1.0, -1.0, -1.0};  // widths 16, 32, 64; -1 where not a candidate
    };
    const std::vector<TuneRecord>& tuning() const { return tuning_; }
    // Sets the back-transformation width of every bucket where it is admissible
    // (16, 32; 64 up to 256 rows), in place of the tuned one.
    Status set_backtransform_width(int cols) {
        if (cols != 16 && cols != 32 && cols != 64)
            return Status::error(kErrInvalidInput, "back-transformation width must be 16, 32 or 64");
        for (auto& bk : buckets_)
            if (bk.ld <= kEigMaxDim && backtransform_width_ok(bk.ld, cols)) bk.bt_cols = cols;
        return Status{};
    }

   private:
    struct Bucket {
        int ld = 0;
        std::vector<int> inst;            // instance ids
        std::vector<int> sel;             // instances selected for the current run
        bool atoms_all = true;            // `atoms` lists every atom of the bucket
        double* arena = nullptr;          // matrices of the bucket
        hipStream_t stream = nullptr;
        rocblas_handle handle = nullptr;
        // scratch per bucket
        double** ptrs = nullptr;          // device pointer arrays (kPtrRoles * count)
        int* act = nullptr;               // device active list
        double* eig = nullptr;            // count * ld eigenvalues (active order)
        double* eig_e = nullptr;          // syevd off-diagonal workspace
        double* resid = nullptr;
        int* sweeps = nullptr;
        int* info = nullptr;
        double* beta = nullptr;           // Broyden scratch for long histories
        size_t beta_cap = 0;
        double* tau = nullptr;            // Householder factors (count * ld)
        double* lu = nullptr;             // inverse-iteration scratch
        int kcap = 0;                     // eigenvector columns kept per instance
        hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around the tridiagonal reduction
        double ev_bytes = 0.0;            // algorithmic bytes of the timed launch
        bool ev_pending = false;
        std::vector<int> active;
        int kP = 0;
        // independent SCC progress: the states of the active instances are
        // gathered straight into pinned host memory at the end of every
        // iteration, and `done` marks their arrival
        SccState* st_host = nullptr;
        hipEvent_t done = nullptr;
        bool inflight = false;
        bool in_setup = false;  // `done` marks the end of the bucket's setup
        int iter = 0;  // SCC iterations of the current run
        // pinned staging of the pointer tables and the active list (the bucket's
        // stream is idle whenever they are rewritten)
        double** h_ptrs = nullptr;
        int* h_act = nullptr;
        double* h_aux = nullptr;  // level shifts to the device, rocSOLVER info back (count entries)
        // the bucket's atoms and its own pair-gradient sums (separate allocations,
        // so buckets that finish at different times never write the same cache line)
        int* atoms = nullptr;
        int natoms = 0;
        bool has_pbc = false;             // a periodic instance is in the bucket
        double* gpair = nullptr;  // 3 * natoms
        double* dedcn = nullptr;  // natoms
        // the bucket's periodic instances and their atoms (Ewald gradient)
        int* pinst = nullptr;
        int npinst = 0, max_ng = 0;
        int* patoms = nullptr;
        int npatoms = 0;
        // Captured iteration (own eigensolver buckets): one graph per state
        // slot, the two differing only in the pinned buffer the states are
        // gathered into.  The pointer tables list every instance of the bucket
        // at its fixed position; k_gate_active, the graph's first node, marks
        // the positions whose iteration is not a plain one with -1.
        int bt_cols = kBacktransformDefaultCols;  // back-transformation width (autotune)
        int g_bt_cols = 0;                // the width the graphs were captured with
        hipGraphExec_t gexec[2] = {nullptr, nullptr};
        int g_nodes = 0;                  // kernels per replay
        int g_kcap = 0;                   // eigenvector capacity the graphs were captured with
        double* g_lu = nullptr;           // and the inverse-iteration scratch
        double** g_ptrs = nullptr;        // kActRoles * count, every position, parity 0
        int* g_inst = nullptr;            // instance at each position
        int* g_act = nullptr;             // gated active list of the replay
        SccState* g_st[2] = {nullptr, nullptr};  // pinned, count entries each
        hipEvent_t g_ev[2] = {nullptr, nullptr};
        bool graph = false;               // the current run iterates this bucket by replays
        bool g_leave = false;             // a non-plain iteration came up: back to the host loop
        long long g_issued = 0, g_done = 0;  // replays of the current run
    };

    Status setup_phase();
    Status scc_phase();
    Status gradient_phase();
    Status launch_iteration(Bucket& bk);
    Status finish_iteration(Bucket& bk);
    Status launch_bucket_gradient(Bucket& bk);
    Status launch_bucket_setup(Bucket& bk);
    Status launch_integrals(Bucket& bk);
    Status upload_active(Bucket& bk, const std::vector<int>& list, int parity);
    Status grow_kcap(Bucket& bk, int kcap);
    Status eigensolve(Bucket& bk, int count);
    // The engine's own eigensolver (automatic choice, up to kEigMaxDim AO);
    // otherwise the rocSOLVER solver the options name.
    bool own_eigensolver(const Bucket& bk) const { return opt_.eigensolver == 0 && bk.ld <= kEigMaxDim; }
    // Above kEigMaxDim with the automatic choice: the engine's own tridiagonal
    // reduction, rocSOLVER for the tridiagonal eigenproblem and the
    // back-transformation of the kept eigenvectors.
    bool own_reduction(const Bucket& bk) const {
        return opt_.eigensolver == 0 && bk.ld > kEigMaxDim && bk.ld <= kLargeTridiagMaxDim;
    }
    Status large_eigensolve(Bucket& bk);
    bool graph_bucket(const Bucket& bk) const;
    Status capture_iteration(Bucket& bk);
    Status enqueue_plain_iteration(Bucket& bk, int slot);
    Status issue_replay(Bucket& bk);
    Status finish_replay(Bucket& bk, bool& switch_to_host);

    const DeviceModel* dm_ = nullptr;
    EngineOptions opt_;
    int ninst_ = 0, natoms_ = 0, nshells_ = 0, naos_ = 0;
    std::vector<int> h_atom0_, h_sh0_, h_ao0_, h_nsh_, h_nao_, h_ld_, h_atom_z_, h_sh_atom_, h_sh_type_, h_sh_ao_, h_ao_sh_,
        h_atom_sh0_, h_atom_inst_, h_sh_by_l_, h_lofs_;
    std::vector<double> h_x_, h_q0_, h_nelec_;
    std::vector<int64_t> h_vec0_, h_broff_;
    std::vector<int> h_bucket_of_, h_final_parity_;
    std::vector<int> h_bucket_pos_;  // position of each instance in its bucket
    std::vector<int> h_err_, h_err_input_;
    std::vector<unsigned char> h_sel_;    // instances selected for the current run
    std::vector<std::string> h_errmsg_, h_errmsg_input_;
    std::vector<Bucket> buckets_;
    std::vector<void*> allocs_;
    SccParams prm_{};
    int broyden_memory_ = 2;
    double timings_[4] = {0, 0, 0, 0};
    const GuessModel* guess_ = nullptr;
    bool q0_dirty_ = false;                  // d_q0_ holds a learned start vector
    std::vector<int> h_guess_usable_;
    std::vector<double> h_charge_;
    int* d_guess_usable_ = nullptr;
    double* d_charge_ = nullptr;
    float *d_gfeat_ = nullptr, *d_gout_ = nullptr;
    double tri_ms_ = 0.0, tri_bytes_ = 0.0;
    long long tri_launches_ = 0;
    LaunchStats launch_stats_{};
    LargeTridiagWork large_work_;  // scratch of the reduction above kEigMaxDim
    std::vector<TuneRecord> tuning_;
    // periodic instances: cell records, reciprocal vectors, Ewald shell matrices
    std::vector<unsigned char> h_pbc_;
    std::vector<int> h_pinst_, h_patoms_;
    int max_pnat_ = 0;
    CellRec* d_cell_ = nullptr;
    double *d_gtab_ = nullptr, *d_gam_ = nullptr, *d_gsf_ = nullptr;
    double* d_ket_ = nullptr;  // ket-share rows of the periodic pair gradient (nullptr: both sides evaluated)
    double* d_gew_ = nullptr;  // Ewald gradient of the periodic atoms (3 per atom)
    // The Ewald gradient of a bucket's periodic instances runs on ew_stream_,
    // beside the bucket's pair gradient, once its SCC is done (ev_ew_q_); the
    // gradient phase waits for ev_ew_.
    hipStream_t ew_stream_ = nullptr;
    hipEvent_t ev_ew_q_ = nullptr, ev_ew_ = nullptr;
    int *d_pinst_ = nullptr, *d_patoms_ = nullptr;
    bool periodic(int inst) const { return !h_pbc_.empty() && h_pbc_[inst]; }

    // device arrays
    DeviceBatch db_{};
    int *d_atom0_ = nullptr, *d_sh0_ = nullptr, *d_ao0_ = nullptr, *d_ld_ = nullptr, *d_atom_inst_ = nullptr,
        *d_atom_z_ = nullptr, *d_atom_sh0_ = nullptr, *d_sh_atom_ = nullptr, *d_sh_type_ = nullptr,
        *d_sh_ao_ = nullptr, *d_ao_sh_ = nullptr, *d_fail_ = nullptr, *d_sh_by_l_ = nullptr, *d_lofs_ = nullptr,
        *d_nao_ = nullptr, *d_nsh_ = nullptr;
    int64_t *d_vec0_ = nullptr, *d_broff_ = nullptr;
    double *d_x_ = nullptr, *d_cn_ = nullptr, *d_cnd3_ = nullptr, *d_gw_ = nullptr;
    double *d_qin_ = nullptr, *d_qout_ = nullptr, *d_q0_ = nullptr, *d_vin_g_ = nullptr, *d_vin_t_ = nullptr,
           *d_vout_g_ = nullptr, *d_vout_t_ = nullptr;
    double *d_rowS_ = nullptr, *d_rowH_ = nullptr, *d_eps_ = nullptr, *d_occ_ = nullptr;
    double *d_brdf_ = nullptr, *d_bru_ = nullptr, *d_brom_ = nullptr, *d_brql_ = nullptr, *d_brdql_ = nullptr;
    double *d_erep_atom_ = nullptr, *d_edisp_atom_ = nullptr, *d_erep_ = nullptr, *d_edisp_ = nullptr,
           *d_ehal_ = nullptr;
    double *d_grep_ = nullptr, *d_gdisp_ = nullptr, *d_ghal_ = nullptr, *d_gel_ = nullptr, *d_dedcn_disp_ = nullptr,
           *d_dedcn_h_ = nullptr;
    double** d_mat_ = nullptr;  // 9 pointer tables of ninst entries
    SccState* d_state_ = nullptr;
    std::vector<SccState> h_state_;
    hipStream_t stream_ = nullptr;
    bool ran_ = false;
    bool progress_ = false;  // GFN1_PROGRESS: one stderr line per bucket iteration
    // GFN1_SERIAL_BUCKETS (profiling only): every bucket's setup, iteration and
    // gradient runs alone on the device, so kernel durations are those of the
    // kernel by itself rather than shared with the other buckets' work
    bool serial_ = false;
    std::chrono::steady_clock::time_point t_scc_{}, setup_end_{};
    hipEvent_t ev_cn_ = nullptr, ev_start_ = nullptr;  // coordination numbers; start charges and geometry terms
    bool has_d_ = false;  // any d shell in the batch
    int pad_multiple_ = kPadMultiple;

    template <typename T>
    Status dalloc(T** p, size_t count);
};

}  // namespace gfn1
