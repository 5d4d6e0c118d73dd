// Batched GFN1-xTB SCC energy + analytic gradient driver.
//
// Flow of Batch::run (reference: run_electronic electronic.rs:800-2031 and
// analytic_gradient_from_result gradient.rs:122-242):
//   setup     CN, S/H0, Cholesky-based orthonormal start basis, geometry terms
//   SCC       every size bucket iterates on its own stream at its own pace
//             (launched largest first, no barrier between buckets):
//             v(q_in) -> A = Cb^T F Cb (reduced congruence, F not formed)
//             -> A = Q T Q^T -> eigenvalues, Fermi, occupied eigenvectors Z
//             -> C = Cb Q Z -> P -> Mulliken q_out -> E, convergence, mixing;
//             a bucket whose instances are all done goes on to W and its pair
//             gradient while the others still iterate
//   gradient  per-atom scatter of the bucket sums, atom-wise terms, CN chains.
#include "engine.hpp"

#include "cholesky.hpp"
#include "density.hpp"
#include "eigen.hpp"
#include "tridiag_large.hpp"

#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace gfn1 {

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return Status::error(kErrDevice, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define RB_TRY(expr)                                                                                   \
    do {                                                                                               \
        rocblas_status s_ = (expr);                                                                    \
        if (s_ != rocblas_status_success)                                                              \
            return Status::error(kErrDevice, std::string(#expr) + ": " + rocblas_status_to_string(s_)); \
    } while (0)
#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

namespace {
// R_L: Cholesky factor of S; R_A0: Cb^T H0 Cb (both fixed for the run)
enum Role { R_S = 0, R_H0, R_F, R_P, R_W, R_CB0, R_CB1, R_T, R_A, R_L, R_A0, kRoles };
// A_C32: FP32 operand behind the A32 operand in the A matrix (mixed precision)
enum ActRole { A_F = 0, A_CB, A_C, A_T, A_A, A_P, A_S, A_X, A_C32, A_H0, A_L, A_A0, kActRoles };

int round_up(int v, int m) { return (v + m - 1) / m * m; }
// hipMemcpy that skips empty ranges (an empty host vector has no buffer, and a
// null pointer is an invalid argument even for zero bytes)
hipError_t copy_bytes(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    return bytes ? hipMemcpy(dst, src, bytes, kind) : hipSuccess;
}
// Each instance's ranges in the per-shell and per-AO arrays start on a 256-byte
// boundary: the size buckets update their instances on concurrent streams, and
// the XCDs' L2 caches are only made coherent at kernel boundaries, so no cache
// line may hold data of two instances.
constexpr int kRangeAlign = 32;
}  // namespace

// ------------------------------------------------------------------ DeviceModel
Status DeviceModel::init(const Model& m) {
    model_ = m;
    const auto& sh = m.shell_table();
    const auto& el = m.element_table();
    HIP_TRY(hipMalloc(&shells_, sh.size() * sizeof(ShellType)));
    HIP_TRY(hipMemcpy(shells_, sh.data(), sh.size() * sizeof(ShellType), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&elements_, el.size() * sizeof(ElementType)));
    HIP_TRY(hipMemcpy(elements_, el.data(), el.size() * sizeof(ElementType), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c6_, m.d3_c6().size() * sizeof(double)));
    HIP_TRY(hipMemcpy(c6_, m.d3_c6().data(), m.d3_c6().size() * sizeof(double), hipMemcpyHostToDevice));
    std::vector<double> pair((kMaxZ + 1) * (kMaxZ + 1), 1.0);
    for (int a = 1; a <= kMaxZ; ++a)
        for (int b = 1; b <= kMaxZ; ++b) pair[a * (kMaxZ + 1) + b] = m.pair_scaling(a, b);
    HIP_TRY(hipMalloc(&pair_, pair.size() * sizeof(double)));
    HIP_TRY(hipMemcpy(pair_, pair.data(), pair.size() * sizeof(double), hipMemcpyHostToDevice));
    return Status{};
}

DeviceModel::~DeviceModel() {
    (void)hipFree(shells_);
    (void)hipFree(elements_);
    (void)hipFree(c6_);
    (void)hipFree(pair_);
}

DeviceTables DeviceModel::tables() const {
    DeviceTables t;
    t.shells = shells_;
    t.elements = elements_;
    t.d3_c6 = c6_;
    t.pair_scale = pair_;
    t.glob = model_.globals();
    return t;
}

// ------------------------------------------------------------------ Batch
template <typename T>
Status Batch::dalloc(T** p, size_t count) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    HIP_TRY(hipMalloc((void**)p, bytes));
    HIP_TRY(hipMemset(*p, 0, bytes));
    allocs_.push_back(*p);
    return Status{};
}

Batch::~Batch() {
    for (auto& bk : buckets_) {
        if (bk.handle) rocblas_destroy_handle(bk.handle);
        if (bk.ev0) (void)hipEventDestroy(bk.ev0);
        if (bk.ev1) (void)hipEventDestroy(bk.ev1);
        if (bk.stream) (void)hipStreamDestroy(bk.stream);
        (void)hipFree(bk.arena);
        (void)hipFree(bk.ptrs);
        (void)hipFree(bk.act);
        (void)hipFree(bk.eig);
        (void)hipFree(bk.eig_e);
        (void)hipFree(bk.resid);
        (void)hipFree(bk.sweeps);
        (void)hipFree(bk.info);
        (void)hipFree(bk.beta);
        (void)hipFree(bk.tau);
        (void)hipFree(bk.lu);
        if (bk.h_ptrs) (void)hipHostFree(bk.h_ptrs);
        if (bk.h_act) (void)hipHostFree(bk.h_act);
        if (bk.h_aux) (void)hipHostFree(bk.h_aux);
        if (bk.st_host) (void)hipHostFree(bk.st_host);
        if (bk.done) (void)hipEventDestroy(bk.done);
        (void)hipFree(bk.atoms);
        (void)hipFree(bk.gpair);
        (void)hipFree(bk.dedcn);
        for (int slot = 0; slot < 2; ++slot) {
            if (bk.gexec[slot]) (void)hipGraphExecDestroy(bk.gexec[slot]);
            if (bk.g_st[slot]) (void)hipHostFree(bk.g_st[slot]);
            if (bk.g_ev[slot]) (void)hipEventDestroy(bk.g_ev[slot]);
        }
        (void)hipFree(bk.g_ptrs);
        (void)hipFree(bk.g_inst);
        (void)hipFree(bk.g_act);
    }
    large_work_.release();
    for (void* p : allocs_) (void)hipFree(p);
    if (ev_cn_) (void)hipEventDestroy(ev_cn_);
    if (ev_ew_q_) (void)hipEventDestroy(ev_ew_q_);
    if (ev_ew_) (void)hipEventDestroy(ev_ew_);
    if (ew_stream_) (void)hipStreamDestroy(ew_stream_);
    if (ev_start_) (void)hipEventDestroy(ev_start_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

Status Batch::create(const DeviceModel& dm, const std::vector<SystemInput>& systems, const EngineOptions& opt) {
    dm_ = &dm;
    opt_ = opt;
    if (const char* pm = std::getenv("GFN1_PAD_MULTIPLE")) pad_multiple_ = std::max(kPadMultiple, std::atoi(pm));
    const Model& m = dm.model();
    ninst_ = (int)systems.size();
    h_atom0_.assign(ninst_ + 1, 0);
    h_sh0_.assign(ninst_ + 1, 0);
    h_ao0_.assign(ninst_ + 1, 0);
    h_nsh_.assign(ninst_, 0);
    h_nao_.assign(ninst_, 0);
    h_ld_.assign(ninst_, 0);
    h_vec0_.assign(ninst_ + 1, 0);
    h_err_.assign(ninst_, 0);
    h_errmsg_.assign(ninst_, "");
    h_nelec_.assign(ninst_, 0.0);
    h_guess_usable_.assign(ninst_, 1);
    h_charge_.assign(ninst_, 0.0);
    h_pbc_.assign(ninst_, 0);
    std::vector<CellRec> cells(ninst_);
    std::vector<double> gtab;
    int64_t gam_size = 0, ket_size = 0;
    for (int i = 0; i < ninst_; ++i) {
        const SystemInput& s = systems[i];
        h_charge_[i] = s.charge;
        std::memset(&cells[i], 0, sizeof(CellRec));
        int nsh = 0, nao = 0;
        double nref = 0.0;
        for (int a = 0; a < s.nat; ++a) {
            int z = s.z[a];
            if (z < 1 || z > kMaxZ || !m.element(z).present)
                return Status::error(kErrMissingElement, "missing element Z=" + std::to_string(z));
            const ElementType& e = m.element(z);
            if (guess_element_slot(z) < 0 || e.nshell != 2) h_guess_usable_[i] = 0;
            h_atom_inst_.push_back(i);
            h_atom_z_.push_back(z);
            h_atom_sh0_.push_back(h_sh0_[i] + nsh);
            for (int k = 0; k < 3; ++k) {
                double v = s.x[3 * a + k];
                if (!std::isfinite(v)) {
                    h_err_[i] = kErrInvalidInput;
                    h_errmsg_[i] = "non-finite coordinate";
                    v = 0.0;
                }
                h_x_.push_back(v);
            }
            for (int sh = 0; sh < e.nshell; ++sh) {
                const ShellType& t = m.shell(z, sh);
                h_sh_atom_.push_back(h_atom0_[i] + a);
                h_sh_type_.push_back(z * kMaxShellsPerAtom + sh);
                h_sh_ao_.push_back(nao);
                for (int k = 0; k < t.nao; ++k) h_ao_sh_.push_back(h_sh0_[i] + nsh);
                nao += t.nao;
                nref += t.ref_occ;
                ++nsh;
            }
        }
        h_atom0_[i + 1] = h_atom0_[i] + s.nat;
        h_nsh_[i] = nsh;
        h_nao_[i] = nao;
        h_sh0_[i + 1] = round_up(h_sh0_[i] + nsh, kRangeAlign);
        h_ao0_[i + 1] = round_up(h_ao0_[i] + nao, kRangeAlign);
        h_ld_[i] = std::max(pad_multiple_, round_up(nao, pad_multiple_));
        h_vec0_[i + 1] = h_vec0_[i] + h_ld_[i];
        const double nelec = nref - s.charge;
        h_nelec_[i] = nelec;
        if (nao == 0) {
            h_err_[i] = kErrInvalidInput;
            h_errmsg_[i] = "cannot run GFN1 SCC for an empty basis";
        } else if (!std::isfinite(nelec) || nelec < -1e-8 || nelec > 2.0 * nao + 1e-8) {
            h_err_[i] = kErrInvalidInput;
            h_errmsg_[i] = "electron count out of range";
        }
        if (s.cell) {
            // Lattice::new (lattice.rs:73-87): a singular cell is an input error
            if (!make_cell_record(s.cell, cells[i], gtab)) {
                if (!h_err_[i]) {
                    h_err_[i] = kErrSingularCell;
                    h_errmsg_[i] = "singular cell";
                }
                std::memset(&cells[i], 0, sizeof(CellRec));
            } else {
                h_pbc_[i] = 1;
                cells[i].gam0 = gam_size;
                gam_size += round_up(nsh * nsh, kRangeAlign);
                // 256-byte aligned: buckets on different streams write their
                // instances' rows concurrently
                cells[i].ket0 = ket_size;
                ket_size += (4 * (int64_t)s.nat * s.nat + kRangeAlign - 1) / kRangeAlign * kRangeAlign;
                h_pinst_.push_back(i);
                max_pnat_ = std::max(max_pnat_, s.nat);
                for (int a = 0; a < s.nat; ++a) h_patoms_.push_back(h_atom0_[i] + a);
            }
        }
        // initial shell charges (electronic.rs:1235-1239): used when the length matches
        for (int k = 0; k < nsh; ++k)
            h_q0_.push_back((s.q0 && s.q0_len == nsh) ? s.q0[k] : 0.0);
        // filler entries up to the next aligned start (never part of an instance)
        for (int k = h_sh0_[i] + nsh; k < h_sh0_[i + 1]; ++k) {
            h_sh_atom_.push_back(s.nat > 0 ? h_atom0_[i] : 0);
            h_sh_type_.push_back(nsh > 0 ? h_sh_type_.back() : 1 * kMaxShellsPerAtom);
            h_sh_ao_.push_back(0);
            h_q0_.push_back(0.0);
        }
        for (int k = h_ao0_[i] + nao; k < h_ao0_[i + 1]; ++k) h_ao_sh_.push_back(h_sh0_[i]);
    }
    h_err_input_ = h_err_;
    h_errmsg_input_ = h_errmsg_;
    natoms_ = h_atom0_[ninst_];
    nshells_ = h_sh0_[ninst_];
    // shells of each instance grouped by angular momentum (stable within a group)
    h_sh_by_l_.resize(h_sh0_[ninst_]);
    h_lofs_.assign(4 * (size_t)ninst_, 0);
    for (int i = 0; i < ninst_; ++i) {
        int pos = h_sh0_[i];
        for (int l = 0; l < 3; ++l) {
            h_lofs_[4 * i + l] = pos;
            for (int sh = h_sh0_[i]; sh < h_sh0_[i] + h_nsh_[i]; ++sh)
                if (m.shell_table()[h_sh_type_[sh]].l == l) h_sh_by_l_[pos++] = sh;
        }
        h_lofs_[4 * i + 3] = pos;
        if (h_lofs_[4 * i + 3] > h_lofs_[4 * i + 2]) has_d_ = true;
    }
    naos_ = h_ao0_[ninst_];

    // SCC ladder (electronic.rs:1413-1419)
    const double base_mix = std::min(1.0, std::max(0.01, opt.mixing));
    const int robust = std::max(opt.max_scc, 800);
    prm_.energy_tolerance = opt.energy_tolerance;
    prm_.charge_tolerance = opt.charge_tolerance;
    const int rb[4] = {opt.scc_broyden ? 1 : 0, 0, 0, 1};
    const double rm[4] = {base_mix, 0.20, 0.10, base_mix};
    const double rl[4] = {opt.level_shift, std::max(opt.level_shift, 0.20), std::max(opt.level_shift, 0.10), 0.0};
    const double rt[4] = {opt.electronic_temperature, opt.electronic_temperature, opt.electronic_temperature,
                          std::max(opt.electronic_temperature, 3000.0)};
    const int ri[4] = {opt.max_scc, robust, robust, std::max(opt.max_scc, 250)};
    for (int k = 0; k < 4; ++k) {
        prm_.rung_broyden[k] = rb[k];
        prm_.rung_mixing[k] = rm[k];
        prm_.rung_level_shift[k] = rl[k];
        prm_.rung_temperature[k] = rt[k];
        prm_.rung_max_iter[k] = ri[k];
    }
    broyden_memory_ = std::max(opt.scc_broyden_size, 2);
    int max_broyden_iter = std::max(opt.scc_broyden ? opt.max_scc : 0, std::max(opt.max_scc, 250));
    prm_.broyden_memory = broyden_memory_;
    prm_.broyden_slots = std::min(broyden_memory_, max_broyden_iter);
    h_broff_.assign(ninst_ + 1, 0);
    for (int i = 0; i < ninst_; ++i)
        h_broff_[i + 1] = h_broff_[i] + round_up(prm_.broyden_slots * h_nsh_[i], kRangeAlign);
    prm_.omega_stride = round_up(prm_.broyden_slots, kRangeAlign);

    HIP_TRY(hipStreamCreate(&stream_));
    HIP_TRY(hipEventCreateWithFlags(&ev_cn_, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&ev_start_, hipEventDisableTiming));
    // ---- device index tables
    auto up_i = [&](int** d, const std::vector<int>& h) -> Status {
        Status st = dalloc(d, h.size());
        if (!st.ok()) return st;
        HIP_TRY(copy_bytes(*d, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
        return Status{};
    };
    Status st;
    if (!(st = up_i(&d_atom0_, h_atom0_)).ok()) return st;
    if (!(st = up_i(&d_sh0_, h_sh0_)).ok()) return st;
    if (!(st = up_i(&d_ao0_, h_ao0_)).ok()) return st;
    if (!(st = up_i(&d_ld_, h_ld_)).ok()) return st;
    if (!(st = up_i(&d_atom_inst_, h_atom_inst_)).ok()) return st;
    if (!(st = up_i(&d_atom_z_, h_atom_z_)).ok()) return st;
    if (!(st = up_i(&d_atom_sh0_, h_atom_sh0_)).ok()) return st;
    if (!(st = up_i(&d_sh_atom_, h_sh_atom_)).ok()) return st;
    if (!(st = up_i(&d_sh_type_, h_sh_type_)).ok()) return st;
    if (!(st = up_i(&d_sh_ao_, h_sh_ao_)).ok()) return st;
    if (!(st = up_i(&d_ao_sh_, h_ao_sh_)).ok()) return st;
    if (!(st = up_i(&d_sh_by_l_, h_sh_by_l_)).ok()) return st;
    if (!(st = up_i(&d_lofs_, h_lofs_)).ok()) return st;
    if (!(st = up_i(&d_nao_, h_nao_)).ok()) return st;
    if (!(st = up_i(&d_nsh_, h_nsh_)).ok()) return st;
    if (!(st = dalloc(&d_fail_, ninst_)).ok()) return st;
    if (!(st = dalloc(&d_vec0_, ninst_ + 1)).ok()) return st;
    HIP_TRY(copy_bytes(d_vec0_, h_vec0_.data(), (ninst_ + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    if (!(st = dalloc(&d_broff_, ninst_ + 1)).ok()) return st;
    HIP_TRY(copy_bytes(d_broff_, h_broff_.data(), (ninst_ + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    if (!(st = dalloc(&d_x_, h_x_.size())).ok()) return st;
    HIP_TRY(copy_bytes(d_x_, h_x_.data(), h_x_.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!(st = dalloc(&d_q0_, nshells_)).ok()) return st;
    HIP_TRY(copy_bytes(d_q0_, h_q0_.data(), nshells_ * sizeof(double), hipMemcpyHostToDevice));
    // Mulliken partial sums: one slice of naos_ per 64-column block of the largest matrix
    int max_ld = 0;
    for (int i = 0; i < ninst_; ++i) max_ld = std::max(max_ld, h_ld_[i]);
    const size_t rowsum_blocks = (size_t)std::max(1, density_row_blocks(max_ld));
    double** dd[] = {&d_cn_, &d_cnd3_, &d_qin_, &d_qout_, &d_vin_g_, &d_vin_t_, &d_vout_g_, &d_vout_t_,
                     &d_rowS_, &d_rowH_, &d_erep_atom_, &d_edisp_atom_, &d_erep_, &d_edisp_, &d_ehal_,
                     &d_grep_, &d_gdisp_, &d_ghal_, &d_gel_, &d_dedcn_disp_, &d_dedcn_h_, &d_brql_, &d_brdql_};
    size_t dn[] = {(size_t)natoms_, (size_t)natoms_, (size_t)nshells_, (size_t)nshells_, (size_t)nshells_,
                   (size_t)nshells_, (size_t)nshells_, (size_t)nshells_, rowsum_blocks * (size_t)naos_,
                   rowsum_blocks * (size_t)naos_,
                   (size_t)natoms_, (size_t)natoms_, (size_t)ninst_, (size_t)ninst_, (size_t)ninst_,
                   3 * (size_t)natoms_, 3 * (size_t)natoms_, 3 * (size_t)natoms_, 3 * (size_t)natoms_,
                   (size_t)natoms_, (size_t)natoms_, (size_t)nshells_, (size_t)nshells_};
    for (size_t k = 0; k < sizeof(dn) / sizeof(dn[0]); ++k)
        if (!(st = dalloc(dd[k], dn[k])).ok()) return st;
    if (!(st = dalloc(&d_gw_, 2 * (size_t)natoms_ * kMaxRef)).ok()) return st;
    if (!(st = dalloc(&d_eps_, h_vec0_[ninst_])).ok()) return st;
    if (!(st = dalloc(&d_occ_, h_vec0_[ninst_])).ok()) return st;
    if (!(st = dalloc(&d_brdf_, h_broff_[ninst_])).ok()) return st;
    if (!(st = dalloc(&d_bru_, h_broff_[ninst_])).ok()) return st;
    if (!(st = dalloc(&d_brom_, (size_t)ninst_ * prm_.omega_stride)).ok()) return st;
    if (!(st = dalloc(&d_state_, ninst_)).ok()) return st;
    if (!(st = dalloc(&d_mat_, (size_t)kRoles * ninst_)).ok()) return st;
    if (!h_pinst_.empty()) {
        if (!(st = dalloc(&d_cell_, ninst_)).ok()) return st;
        HIP_TRY(copy_bytes(d_cell_, cells.data(), ninst_ * sizeof(CellRec), hipMemcpyHostToDevice));
        if (!(st = dalloc(&d_gtab_, std::max<size_t>(gtab.size(), 1))).ok()) return st;
        HIP_TRY(copy_bytes(d_gtab_, gtab.data(), gtab.size() * sizeof(double), hipMemcpyHostToDevice));
        if (!(st = dalloc(&d_gam_, std::max<int64_t>(gam_size, 1))).ok()) return st;
        if (!(st = dalloc(&d_gsf_, gtab.size() / kGStride * kGsfStride)).ok()) return st;
        // GFN1_KET_SHARE_MAX_ATOMS lowers the limit (0: every pair from both sides)
        int ket_max = kKetShareMaxAtoms;
        if (const char* km = std::getenv("GFN1_KET_SHARE_MAX_ATOMS"))
            ket_max = std::min(kKetShareMaxAtoms, std::max(0, std::atoi(km)));
        if (max_pnat_ <= ket_max && !(st = dalloc(&d_ket_, ket_size)).ok()) return st;
        if (!(st = dalloc(&d_gew_, 3 * (size_t)natoms_)).ok()) return st;
        HIP_TRY(hipStreamCreate(&ew_stream_));
        HIP_TRY(hipEventCreateWithFlags(&ev_ew_q_, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&ev_ew_, hipEventDisableTiming));
        if (!(st = up_i(&d_pinst_, h_pinst_)).ok()) return st;
        if (!(st = up_i(&d_patoms_, h_patoms_)).ok()) return st;
    }

    // ---- size buckets and matrix arenas
    std::vector<int> lds = h_ld_;
    std::sort(lds.begin(), lds.end());
    lds.erase(std::unique(lds.begin(), lds.end()), lds.end());
    h_bucket_of_.assign(ninst_, 0);
    h_bucket_pos_.assign(ninst_, 0);
    buckets_.resize(lds.size());
    std::vector<double*> hmat((size_t)kRoles * ninst_, nullptr);
    for (size_t bi = 0; bi < lds.size(); ++bi) buckets_[bi].ld = lds[bi];
    for (int i = 0; i < ninst_; ++i) {
        int bi = (int)(std::lower_bound(lds.begin(), lds.end(), h_ld_[i]) - lds.begin());
        h_bucket_of_[i] = bi;
        h_bucket_pos_[i] = (int)buckets_[bi].inst.size();
        buckets_[bi].inst.push_back(i);
    }
    for (auto& bk : buckets_) {
        const size_t mat = (size_t)bk.ld * bk.ld;
        const size_t count = bk.inst.size();
        HIP_TRY(hipMalloc(&bk.arena, mat * count * kRoles * sizeof(double)));
        HIP_TRY(hipMemset(bk.arena, 0, mat * count * kRoles * sizeof(double)));
        for (size_t k = 0; k < count; ++k)
            for (int r = 0; r < kRoles; ++r)
                hmat[(size_t)r * ninst_ + bk.inst[k]] = bk.arena + ((size_t)r * count + k) * mat;
        HIP_TRY(hipStreamCreate(&bk.stream));
        HIP_TRY(hipEventCreate(&bk.ev0));
        HIP_TRY(hipEventCreate(&bk.ev1));
        RB_TRY(rocblas_create_handle(&bk.handle));
        RB_TRY(rocblas_set_stream(bk.handle, bk.stream));
        // reproducible GEMMs: no split-K kernels that accumulate with atomics
        RB_TRY(rocblas_set_atomics_mode(bk.handle, rocblas_atomics_not_allowed));
        HIP_TRY(hipMalloc(&bk.ptrs, sizeof(double*) * kActRoles * count));
        HIP_TRY(hipMalloc(&bk.act, sizeof(int) * count));
        HIP_TRY(hipMalloc(&bk.eig, sizeof(double) * bk.ld * count));
        HIP_TRY(hipMalloc(&bk.eig_e, sizeof(double) * bk.ld * count));
        HIP_TRY(hipMalloc(&bk.resid, sizeof(double) * count));
        HIP_TRY(hipMalloc(&bk.sweeps, sizeof(int) * count));
        HIP_TRY(hipMalloc(&bk.info, sizeof(int) * count));
        HIP_TRY(hipMalloc(&bk.tau, sizeof(double) * bk.ld * count));
        // eigenvector capacity: occupied orbitals of the most electron-rich
        // instance plus a margin for fractional occupations
        int kmax = 1;
        for (int i : bk.inst) kmax = std::max(kmax, (int)std::ceil(0.5 * h_nelec_[i] - 1e-9));
        bk.kcap = opt.electronic_temperature > kWideSmearing ? bk.ld : std::min(bk.ld, kmax + 32);
        HIP_TRY(hipMalloc(&bk.lu, sizeof(double) * 5 * (size_t)bk.ld * bk.kcap * count));
        // Per-iteration host traffic goes through coherent pinned buffers that
        // small kernels read and write directly: an asynchronous copy would go
        // to a copy engine shared by all streams, where a copy still waiting for
        // its own stream's kernels holds up the copies of every other bucket.
        HIP_TRY(hipHostMalloc((void**)&bk.h_ptrs, sizeof(double*) * kActRoles * count, hipHostMallocCoherent));
        HIP_TRY(hipHostMalloc((void**)&bk.h_act, sizeof(int) * count, hipHostMallocCoherent));
        HIP_TRY(hipHostMalloc((void**)&bk.h_aux, sizeof(double) * count, hipHostMallocCoherent));
        HIP_TRY(hipHostMalloc((void**)&bk.st_host, sizeof(SccState) * count, hipHostMallocCoherent));
        HIP_TRY(hipEventCreateWithFlags(&bk.done, hipEventDisableTiming));
        HIP_TRY(hipMalloc(&bk.g_ptrs, sizeof(double*) * kActRoles * count));
        HIP_TRY(hipMalloc(&bk.g_inst, sizeof(int) * count));
        HIP_TRY(hipMalloc(&bk.g_act, sizeof(int) * count));
        for (int slot = 0; slot < 2; ++slot) {
            HIP_TRY(hipHostMalloc((void**)&bk.g_st[slot], sizeof(SccState) * count, hipHostMallocCoherent));
            HIP_TRY(hipEventCreateWithFlags(&bk.g_ev[slot], hipEventDisableTiming));
        }
        std::vector<int> atoms, pinst, patoms;
        for (int i : bk.inst) {
            for (int a = h_atom0_[i]; a < h_atom0_[i + 1]; ++a) atoms.push_back(a);
            if (h_pbc_[i]) {
                bk.has_pbc = true;
                pinst.push_back(i);
                bk.max_ng = std::max(bk.max_ng, (int)cells[i].ng);
                for (int a = h_atom0_[i]; a < h_atom0_[i + 1]; ++a) patoms.push_back(a);
            }
        }
        if (!pinst.empty()) {
            Status st;
            if (!(st = up_i(&bk.pinst, pinst)).ok()) return st;
            if (!(st = up_i(&bk.patoms, patoms)).ok()) return st;
            bk.npinst = (int)pinst.size();
            bk.npatoms = (int)patoms.size();
        }
        bk.natoms = (int)atoms.size();
        const size_t na_alloc = std::max<size_t>(atoms.size(), 1);
        HIP_TRY(hipMalloc(&bk.atoms, sizeof(int) * na_alloc));
        if (!atoms.empty())
            HIP_TRY(copy_bytes(bk.atoms, atoms.data(), atoms.size() * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc(&bk.gpair, sizeof(double) * 3 * na_alloc));
        HIP_TRY(hipMalloc(&bk.dedcn, sizeof(double) * na_alloc));
    }
    HIP_TRY(copy_bytes(d_mat_, hmat.data(), hmat.size() * sizeof(double*), hipMemcpyHostToDevice));

    db_.ninst = ninst_;
    db_.atom0 = d_atom0_;
    db_.sh0 = d_sh0_;
    db_.ao0 = d_ao0_;
    db_.ld = d_ld_;
    db_.nsh = d_nsh_;
    db_.nao = d_nao_;
    db_.atom_inst = d_atom_inst_;
    db_.atom_z = d_atom_z_;
    db_.atom_x = d_x_;
    db_.atom_sh0 = d_atom_sh0_;
    db_.sh_atom = d_sh_atom_;
    db_.sh_type = d_sh_type_;
    db_.sh_ao = d_sh_ao_;
    db_.ao_sh = d_ao_sh_;
    db_.sh_by_l = d_sh_by_l_;
    db_.lofs = d_lofs_;
    db_.vec0 = d_vec0_;
    db_.S = d_mat_ + (size_t)R_S * ninst_;
    db_.H0 = d_mat_ + (size_t)R_H0 * ninst_;
    db_.F = d_mat_ + (size_t)R_F * ninst_;
    db_.P = d_mat_ + (size_t)R_P * ninst_;
    db_.W = d_mat_ + (size_t)R_W * ninst_;
    db_.CB0 = d_mat_ + (size_t)R_CB0 * ninst_;
    db_.CB1 = d_mat_ + (size_t)R_CB1 * ninst_;
    db_.T = d_mat_ + (size_t)R_T * ninst_;
    db_.A = d_mat_ + (size_t)R_A * ninst_;
    db_.cell = d_cell_;
    db_.gtab = d_gtab_;
    db_.gam = d_gam_;
    db_.gsf = d_gsf_;
    h_final_parity_.assign(ninst_, 0);
    return Status{};
}

Status Batch::grow_kcap(Bucket& bk, int kcap) {
    HIP_TRY(hipDeviceSynchronize());
    (void)hipFree(bk.lu);
    bk.lu = nullptr;
    HIP_TRY(hipMalloc(&bk.lu, sizeof(double) * 5 * (size_t)bk.ld * kcap * bk.inst.size()));
    bk.kcap = kcap;
    return Status{};
}

// Pointer tables and active list of the listed instances of a bucket, staged in
// pinned host memory and copied by a kernel on the bucket's stream.  The caller
// guarantees that the stream has no pending read of the staging buffers (the
// bucket is idle).
Status Batch::upload_active(Bucket& bk, const std::vector<int>& list, int parity) {
    const size_t mat = (size_t)bk.ld * bk.ld;
    const size_t count = bk.inst.size();
    const size_t m = list.size();
    if (m == 0) return Status{};
    double** hp = bk.h_ptrs;
    for (size_t j = 0; j < m; ++j) {
        const int i = list[j];
        const size_t k = (size_t)h_bucket_pos_[i];
        auto at = [&](int role) { return bk.arena + ((size_t)role * count + k) * mat; };
        const int par = parity < 0 ? h_final_parity_[i] : parity;
        hp[(size_t)A_F * m + j] = at(R_F);
        hp[(size_t)A_CB * m + j] = at(par ? R_CB1 : R_CB0);
        hp[(size_t)A_C * m + j] = at(par ? R_CB0 : R_CB1);
        hp[(size_t)A_T * m + j] = at(R_T);
        hp[(size_t)A_A * m + j] = at(R_A);
        hp[(size_t)A_P * m + j] = at(R_P);
        hp[(size_t)A_S * m + j] = at(R_S);
        hp[(size_t)A_X * m + j] = at(R_W);
        hp[(size_t)A_C32 * m + j] = (double*)((float*)at(R_A) + (size_t)bk.ld * bk.kcap);
        hp[(size_t)A_H0 * m + j] = at(R_H0);
        hp[(size_t)A_L * m + j] = at(R_L);
        hp[(size_t)A_A0 * m + j] = at(R_A0);
        bk.h_act[j] = i;
    }
    const int wp = (int)(kActRoles * m * sizeof(double*) / sizeof(unsigned int));
    hipLaunchKernelGGL(k_copy_words, dim3((wp + 255) / 256), dim3(256), 0, bk.stream, (const unsigned int*)hp,
                       (unsigned int*)bk.ptrs, wp);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_copy_words, dim3(((int)m + 255) / 256), dim3(256), 0, bk.stream,
                       (const unsigned int*)bk.h_act, (unsigned int*)bk.act, (int)m);
    LAUNCH_CHECK();
    return Status{};
}

// Eigenpairs of the active instances of a bucket above kEigMaxDim: the engine's
// own Householder reduction (tridiag_large.hip), rocsolver_dstedc for the
// tridiagonal eigenproblem (eigenvalues ascending into bk.eig, eigenvectors of
// T into the C slot), and rocsolver_dormtr to take the first kcap of them back
// to the basis of A.  Same contract as eigensolve(): eigenvalues at bk.eig +
// l n, convergence flags at bk.info + l; the eigenvectors are in the A_C slot.
Status Batch::large_eigensolve(Bucket& bk) {
    const int n = bk.ld, cnt = (int)bk.active.size();
    const size_t mat = (size_t)n * n, count = bk.inst.size();
    HIP_TRY(large_work_.allocate(n));
    for (int l = 0; l < cnt; ++l) {
        const size_t k = (size_t)h_bucket_pos_[bk.active[l]];
        double* A = bk.arena + ((size_t)R_A * count + k) * mat;
        double* C = bk.arena + ((size_t)R_CB1 * count + k) * mat;  // the A_C slot at parity 0
        double* d = bk.eig + (size_t)l * n;
        double* e = bk.eig_e + (size_t)l * n;
        double* tau = bk.tau + (size_t)l * n;
        HIP_TRY(launch_tridiag_large(bk.stream, n, n, A, d, e, tau, large_work_));
        RB_TRY(rocsolver_dstedc(bk.handle, rocblas_evect_tridiagonal, n, d, e, C, n, bk.info + l));
        RB_TRY(rocsolver_dormtr(bk.handle, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, n, bk.kcap,
                                A, n, tau, C, n));
    }
    return Status{};
}

Status Batch::eigensolve(Bucket& bk, int count) {
    double* const* A = bk.ptrs + (size_t)A_A * count;
    const int n = bk.ld;
    int solver = opt_.eigensolver;
    if (solver == 0) solver = 1;
    if (solver == 3) {
        RB_TRY(rocsolver_dsyevdj_batched(bk.handle, rocblas_evect_original, rocblas_fill_lower, n,
                                         (double* const*)A, n, bk.eig, n, bk.info, count));
    } else if (solver == 2) {
        RB_TRY(rocsolver_dsyevj_batched(bk.handle, rocblas_esort_ascending, rocblas_evect_original,
                                        rocblas_fill_lower, n, (double* const*)A, n, 1e-12, bk.resid, 100,
                                        bk.sweeps, bk.eig, n, bk.info, count));
    } else {
        RB_TRY(rocsolver_dsyevd_batched(bk.handle, rocblas_evect_original, rocblas_fill_lower, n,
                                        (double* const*)A, n, bk.eig, n, bk.eig_e, n, bk.info, count));
    }
    return Status{};
}

Status Batch::run(const unsigned char* select) {
    // an error that an earlier failed call left in this thread's HIP state must
    // not be reported by the launch checks of this run
    (void)hipGetLastError();
    // re-runnable: every run starts from the validated inputs
    h_err_ = h_err_input_;
    h_errmsg_ = h_errmsg_input_;
    // the selected instances of each bucket and their atoms (the atom lists
    // are rewritten only when the selection changes them)
    h_sel_.assign(ninst_, 1);
    if (select)
        for (int i = 0; i < ninst_; ++i) h_sel_[i] = select[i] ? 1 : 0;
    for (auto& bk : buckets_) {
        bk.sel.clear();
        for (int i : bk.inst)
            if (h_sel_[i]) bk.sel.push_back(i);
        const bool all = bk.sel.size() == bk.inst.size();
        if (all && bk.atoms_all) continue;
        std::vector<int> atoms;
        for (int i : bk.sel)
            for (int a = h_atom0_[i]; a < h_atom0_[i + 1]; ++a) atoms.push_back(a);
        bk.natoms = (int)atoms.size();
        if (!atoms.empty())
            HIP_TRY(copy_bytes(bk.atoms, atoms.data(), atoms.size() * sizeof(int), hipMemcpyHostToDevice));
        bk.atoms_all = all;
    }
    if (natoms_ > 0) HIP_TRY(hipMemset(d_ghal_, 0, 3 * (size_t)natoms_ * sizeof(double)));
    HIP_TRY(hipMemset(d_fail_, 0, (size_t)std::max(ninst_, 1) * sizeof(int)));
    ran_ = true;
    serial_ = std::getenv("GFN1_SERIAL_BUCKETS") != nullptr;
    tri_ms_ = tri_bytes_ = 0.0;
    tri_launches_ = 0;
    launch_stats_ = LaunchStats{};
    auto t0 = std::chrono::steady_clock::now();
    setup_end_ = t0;
    // the setup of the buckets overlaps the first SCC iterations of the buckets
    // that finish it first; the phase boundary is the end of the last setup
    Status st = setup_phase();
    if (!st.ok()) return st;
    st = scc_phase();
    if (!st.ok()) return st;
    HIP_TRY(hipDeviceSynchronize());
    auto t1 = setup_end_;
    auto t2 = std::chrono::steady_clock::now();
    st = gradient_phase();
    if (!st.ok()) return st;
    HIP_TRY(hipDeviceSynchronize());
    auto t3 = std::chrono::steady_clock::now();
    auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    timings_[0] = ms(t0, t1);
    timings_[1] = ms(t1, t2);
    timings_[2] = ms(t2, t3);
    timings_[3] = ms(t0, t3);
    return Status{};
}

Status Batch::set_guess(const GuessModel* model) {
    if (model && !d_gfeat_ && natoms_ > 0) {
        Status st;
        if (!(st = dalloc(&d_charge_, ninst_)).ok()) return st;
        HIP_TRY(hipMalloc(&d_guess_usable_, ninst_ * sizeof(int)));
        allocs_.push_back(d_guess_usable_);
        HIP_TRY(hipMalloc(&d_gfeat_, (size_t)natoms_ * kGuessFeat * sizeof(float)));
        allocs_.push_back(d_gfeat_);
        HIP_TRY(hipMalloc(&d_gout_, (size_t)natoms_ * kGuessOut * sizeof(float)));
        allocs_.push_back(d_gout_);
        HIP_TRY(copy_bytes(d_charge_, h_charge_.data(), ninst_ * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(copy_bytes(d_guess_usable_, h_guess_usable_.data(), ninst_ * sizeof(int), hipMemcpyHostToDevice));
    }
    guess_ = model;
    return Status{};
}

Status Batch::set_positions(const double* x) {
    if (!x && natoms_ > 0) return Status::error(kErrInvalidInput, "null positions");
    for (size_t k = 0; k < h_x_.size(); ++k)
        if (!std::isfinite(x[k])) return Status::error(kErrInvalidInput, "non-finite coordinate");
    HIP_TRY(hipDeviceSynchronize());
    std::copy(x, x + h_x_.size(), h_x_.begin());
    if (!h_x_.empty())
        HIP_TRY(copy_bytes(d_x_, h_x_.data(), h_x_.size() * sizeof(double), hipMemcpyHostToDevice));
    // finite coordinates clear the input errors that non-finite ones raised
    for (int i = 0; i < ninst_; ++i)
        if (h_err_input_[i] == kErrInvalidInput && h_errmsg_input_[i] == "non-finite coordinate") {
            h_err_input_[i] = kOk;
            h_errmsg_input_[i].clear();
        }
    return Status{};
}

Status Batch::set_start_charges(int inst, const double* q) {
    if (inst < 0 || inst >= ninst_ || (!q && h_nsh_[inst] > 0))
        return Status::error(kErrInvalidInput, "invalid start charges");
    std::copy(q, q + h_nsh_[inst], h_q0_.begin() + h_sh0_[inst]);
    q0_dirty_ = true;
    return Status{};
}

Status Batch::fetch_energy_gradient(std::vector<int>& status, std::vector<double>& energy,
                                    std::vector<double>& gradient, std::vector<double>* shell_charges,
                                    std::vector<double>* term_gradients) {
    const int na = natoms_;
    std::vector<double> gel(3 * na), grep(3 * na), gdisp(3 * na), ghal(3 * na), erep(ninst_), edisp(ninst_),
        ehal(ninst_);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(copy_bytes(h_state_.data(), d_state_, ninst_ * sizeof(SccState), hipMemcpyDeviceToHost));
    if (na > 0) {
        HIP_TRY(copy_bytes(gel.data(), d_gel_, 3 * na * sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(copy_bytes(grep.data(), d_grep_, 3 * na * sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(copy_bytes(gdisp.data(), d_gdisp_, 3 * na * sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(copy_bytes(ghal.data(), d_ghal_, 3 * na * sizeof(double), hipMemcpyDeviceToHost));
    }
    HIP_TRY(copy_bytes(erep.data(), d_erep_, ninst_ * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(edisp.data(), d_edisp_, ninst_ * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(ehal.data(), d_ehal_, ninst_ * sizeof(double), hipMemcpyDeviceToHost));
    if (shell_charges) {
        shell_charges->assign(nshells_, 0.0);
        if (nshells_ > 0)
            HIP_TRY(copy_bytes(shell_charges->data(), d_qout_, nshells_ * sizeof(double), hipMemcpyDeviceToHost));
    }
    status.assign(ninst_, 0);
    energy.assign(ninst_, 0.0);
    gradient.assign(3 * (size_t)na, 0.0);
    if (term_gradients) term_gradients->assign(9 * (size_t)na, 0.0);
    const bool el_on = opt_.compute_gradient && (opt_.include_hamiltonian || opt_.include_scc);
    const bool rep_on = opt_.compute_gradient && opt_.include_repulsion;
    const bool disp_on = opt_.compute_gradient && opt_.include_dispersion && opt_.enable_dispersion;
    const bool hal_on = opt_.compute_gradient && opt_.include_halogen;
    for (int i = 0; i < ninst_; ++i) {
        const SccState& s = h_state_[i];
        int st = h_err_[i] ? h_err_[i] : (s.status == 1 ? kOk : kErrSccNotConverged);
        if (s.status == 3) st = kErrInvalidInput;
        status[i] = st;
        if (st != kOk) continue;
        const double e_disp = opt_.enable_dispersion ? edisp[i] : 0.0;
        energy[i] = s.e_el + s.e2 + s.e3 + erep[i] + e_disp + ehal[i] + s.ts;
        const bool all = periodic(i) && opt_.compute_gradient;  // pbc/gradient.rs:75-163: every term
        for (int k = 3 * h_atom0_[i]; k < 3 * h_atom0_[i + 1]; ++k) {
            const double r = rep_on || all ? grep[k] : 0.0;
            const double d = disp_on || (all && opt_.enable_dispersion) ? gdisp[k] : 0.0;
            const double h = hal_on || all ? ghal[k] : 0.0;
            gradient[k] = (el_on || all ? gel[k] : 0.0) + r + d + h;
            if (term_gradients) {
                (*term_gradients)[k] = r;
                (*term_gradients)[3 * (size_t)na + k] = d;
                (*term_gradients)[6 * (size_t)na + k] = h;
            }
        }
    }
    return Status{};
}

Status Batch::fetch_dipoles(std::vector<double>& dipoles) {
    std::vector<double> qout(nshells_);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(copy_bytes(qout.data(), d_qout_, nshells_ * sizeof(double), hipMemcpyDeviceToHost));
    dipoles.assign(3 * (size_t)ninst_, 0.0);
    std::vector<double> qa;
    for (int i = 0; i < ninst_; ++i) {
        const int a0 = h_atom0_[i], nat = h_atom0_[i + 1] - a0;
        qa.assign(nat, 0.0);
        for (int k = h_sh0_[i]; k < h_sh0_[i] + h_nsh_[i]; ++k) qa[h_sh_atom_[k] - a0] += qout[k];
        for (int d = 0; d < 3; ++d) {
            double v = 0.0;
            for (int a = 0; a < nat; ++a) v += h_x_[3 * (a0 + a) + d] * qa[a];
            dipoles[3 * (size_t)i + d] = v;
        }
    }
    return Status{};
}

// Setup.  The coordination numbers come first (stream_); then every bucket, on
// its own stream, builds S and H0 over its own atoms, factorises S = L L^T,
// forms Cb = L^-T and A0 = Cb^T H0 Cb, marks instances whose overlap is not
// positive definite, and copies those flags to the host behind its `done` event.
// The geometry terms and the start charges run on stream_ meanwhile; each
// bucket waits for them only before its first SCC iteration.  Buckets above
// kCholMaxDim (single large systems) are set up first, one at a time, with
// rocSOLVER.
Status Batch::setup_phase() {
    const DeviceTables t = dm_->tables();
    const int na = natoms_;
    for (auto& bk : buckets_) bk.in_setup = false;
    // SCC state (nothing runs on the device yet)
    h_state_.assign(ninst_, SccState{});
    for (int i = 0; i < ninst_; ++i) {
        SccState& s = h_state_[i];
        std::memset(&s, 0, sizeof(s));
        s.status = !h_sel_[i] ? 3 : (h_err_[i] ? 2 : 0);
        s.mixing = prm_.rung_mixing[0];
        // the periodic SCC applies no level shift (pbc/scf.rs:358-529)
        s.level_shift = periodic(i) ? 0.0 : prm_.rung_level_shift[0];
        s.temperature = prm_.rung_temperature[0];
        s.kt = std::max(s.temperature, 0.0) * kBoltzmannHartree;
        s.nelec = h_nelec_[i];
    }
    HIP_TRY(copy_bytes(d_state_, h_state_.data(), ninst_ * sizeof(SccState), hipMemcpyHostToDevice));
    if (na == 0) return Status{};

    hipLaunchKernelGGL(k_coordination, dim3((na + 127) / 128), dim3(128), 0, stream_, db_, t,
                       opt_.coordination_cutoff, opt_.enable_cn_hamiltonian, d_cn_, d_cnd3_, d_gw_,
                       d_gw_ + (size_t)na * kMaxRef, na);
    LAUNCH_CHECK();
    if (!h_patoms_.empty()) {
        hipLaunchKernelGGL(k_coordination_pbc, dim3((int)h_patoms_.size()), dim3(64), 0, stream_, db_, t,
                           (const int*)d_patoms_, opt_.coordination_cutoff, opt_.enable_cn_hamiltonian, d_cn_,
                           d_cnd3_, d_gw_, d_gw_ + (size_t)na * kMaxRef);
        LAUNCH_CHECK();
    }
    HIP_TRY(hipEventRecord(ev_cn_, stream_));
    // Ewald shell Coulomb matrices of the periodic instances (needed from the
    // first SCC iteration on; the buckets wait for ev_start_ below)
    HIP_TRY(launch_ewald_gamma(stream_, db_, t, d_pinst_, (int)h_pinst_.size(), max_pnat_));

    const double one = 1.0, zero = 0.0;
    auto a0_gemms = [&](Bucket& bk, int count) -> Status {
        double* const* Cb = bk.ptrs + (size_t)A_CB * count;
        double* const* H0 = bk.ptrs + (size_t)A_H0 * count;
        double* const* T = bk.ptrs + (size_t)A_T * count;
        double* const* A0 = bk.ptrs + (size_t)A_A0 * count;
        const int n = bk.ld;
        RB_TRY(rocblas_dgemm_batched(bk.handle, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                                     (const double* const*)H0, n, (const double* const*)Cb, n, &zero,
                                     (double* const*)T, n, count));
        RB_TRY(rocblas_dgemm_batched(bk.handle, rocblas_operation_transpose, rocblas_operation_none, n, n, n, &one,
                                     (const double* const*)Cb, n, (const double* const*)T, n, &zero,
                                     (double* const*)A0, n, count));
        return Status{};
    };
    // S, H0 and the padding diagonal of a bucket's instances, on its stream
    auto integrals = [&](Bucket& bk, int count) -> Status {
        HIP_TRY(hipStreamWaitEvent(bk.stream, ev_cn_, 0));
        Status st = upload_active(bk, bk.sel, 0);
        if (!st.ok()) return st;
        if (bk.natoms > 0) {
            Status st2 = launch_integrals(bk);
            if (!st2.ok()) return st2;
        }
        hipLaunchKernelGGL(k_pad_diagonal, dim3(count), dim3(64), 0, bk.stream, db_, (const int*)bk.act);
        LAUNCH_CHECK();
        return Status{};
    };

    // Larger matrices (single large systems): rocSOLVER, one bucket at a time,
    // with nothing else on the device.
    bool large = false;
    for (auto& bk : buckets_) {
        if (bk.ld <= kCholMaxDim || bk.sel.empty()) continue;
        large = true;
        const int count = (int)bk.sel.size();
        Status st = integrals(bk, count);
        if (!st.ok()) return st;
        const size_t mat = (size_t)bk.ld * bk.ld;
        const size_t all = bk.inst.size();  // the arena roles hold every instance of the bucket
        // A <- S
        HIP_TRY(hipMemcpyAsync(bk.arena + (size_t)R_A * all * mat, bk.arena + (size_t)R_S * all * mat,
                               all * mat * sizeof(double), hipMemcpyDeviceToDevice, bk.stream));
        double* const* A = bk.ptrs + (size_t)A_A * count;
        RB_TRY(rocsolver_dpotrf_batched(bk.handle, rocblas_fill_lower, bk.ld, (double* const*)A, bk.ld, bk.info,
                                        count));
        // keep L (S = L L^T) for the reduced congruence of the SCC iterations
        HIP_TRY(hipMemcpyAsync(bk.arena + (size_t)R_L * all * mat, bk.arena + (size_t)R_A * all * mat,
                               all * mat * sizeof(double), hipMemcpyDeviceToDevice, bk.stream));
        RB_TRY(rocsolver_dtrtri_batched(bk.handle, rocblas_fill_lower, rocblas_diagonal_non_unit, bk.ld,
                                        (double* const*)A, bk.ld, bk.info, count));
        std::vector<int> info(count);
        HIP_TRY(hipMemcpyAsync(info.data(), bk.info, count * sizeof(int), hipMemcpyDeviceToHost, bk.stream));
        HIP_TRY(hipStreamSynchronize(bk.stream));
        for (int k = 0; k < count; ++k)
            if (info[k] != 0 && !h_err_[bk.sel[k]]) {
                h_err_[bk.sel[k]] = kErrInvalidInput;
                h_errmsg_[bk.sel[k]] = "overlap matrix is not positive definite";
            }
        hipLaunchKernelGGL(k_make_cb, dim3(64, count), dim3(256), 0, bk.stream, db_, A, (const int*)bk.act,
                           opt_.eigen_tolerance, d_fail_);
        LAUNCH_CHECK();
        st = a0_gemms(bk, count);
        if (!st.ok()) return st;
        // With this bucket's batched GEMMs still running on its stream while
        // another bucket's rocSOLVER potrf/trtri ran on another stream, about 1%
        // of the instances of the buckets of 256 AO and more converged to wrong
        // energies (differences up to 1e-2 Eh, different instances on every run,
        // on MI355X / ROCm 7); the rocSOLVER buckets are therefore set up one
        // after another.
        HIP_TRY(hipStreamSynchronize(bk.stream));
    }
    if (large) {
        HIP_TRY(hipDeviceSynchronize());
        std::vector<int> fail(ninst_);
        HIP_TRY(copy_bytes(fail.data(), d_fail_, ninst_ * sizeof(int), hipMemcpyDeviceToHost));
        bool changed = false;
        for (int i = 0; i < ninst_; ++i)
            if ((fail[i] || h_err_[i]) && h_state_[i].status == 0) {
                if (!h_err_[i]) {
                    h_err_[i] = kErrInvalidInput;
                    h_errmsg_[i] = "overlap matrix is not positive definite";
                }
                h_state_[i].status = 2;
                changed = true;
            }
        if (changed)
            HIP_TRY(copy_bytes(d_state_, h_state_.data(), ninst_ * sizeof(SccState), hipMemcpyHostToDevice));
    }

    // Matrices up to kCholMaxDim: every bucket on its own stream, all at once,
    // largest first.
    for (auto bit = buckets_.rbegin(); bit != buckets_.rend(); ++bit) {
        if (bit->ld > kCholMaxDim || bit->sel.empty()) continue;
        Status st = launch_bucket_setup(*bit);
        if (!st.ok()) return st;
        if (serial_) HIP_TRY(hipDeviceSynchronize());
    }

    // geometry-only terms and the SCC start charges (stream_)
    hipLaunchKernelGGL(k_geometry_terms, dim3((na + 127) / 128), dim3(128), 0, stream_, db_, t,
                       (const double*)d_gw_, (const double*)d_cnd3_, opt_.enable_dispersion, d_erep_atom_,
                       d_edisp_atom_, d_grep_, d_gdisp_, d_dedcn_disp_, na);
    LAUNCH_CHECK();
    if (!h_patoms_.empty()) {
        hipLaunchKernelGGL(k_geometry_terms_pbc, dim3((int)h_patoms_.size()), dim3(64), 0, stream_, db_, t,
                           (const int*)d_patoms_, (const double*)d_gw_, opt_.enable_dispersion, d_erep_atom_,
                           d_edisp_atom_, d_grep_, d_gdisp_, d_dedcn_disp_, na);
        LAUNCH_CHECK();
    }
    if (opt_.enable_dispersion) {
        hipLaunchKernelGGL(k_cn_chain, dim3((na + 127) / 128), dim3(128), 0, stream_, db_, t,
                           (const double*)d_dedcn_disp_, 25.0, 1e-14, d_gdisp_, na);
        LAUNCH_CHECK();
        if (!h_patoms_.empty()) {
            hipLaunchKernelGGL(k_cn_chain_pbc, dim3((int)h_patoms_.size()), dim3(64), 0, stream_, db_, t,
                               (const int*)d_patoms_, (const double*)d_dedcn_disp_, 25.0, 1e-14, d_gdisp_);
            LAUNCH_CHECK();
        }
    }
    hipLaunchKernelGGL(k_halogen, dim3((ninst_ + 63) / 64), dim3(64), 0, stream_, db_, t, d_ehal_, d_ghal_);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_instance_sum, dim3(ninst_), dim3(64), 0, stream_, db_, (const double*)d_erep_atom_,
                       d_erep_);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_instance_sum, dim3(ninst_), dim3(64), 0, stream_, db_, (const double*)d_edisp_atom_,
                       d_edisp_);
    LAUNCH_CHECK();
    if (guess_ || q0_dirty_) {
        // the caller's start vector, then the learned one where the model applies
        HIP_TRY(hipMemcpyAsync(d_q0_, h_q0_.data(), nshells_ * sizeof(double), hipMemcpyHostToDevice, stream_));
        q0_dirty_ = false;
    }
    if (guess_) {
        const float *th, *nm;
        Status gs = guess_->device_ptrs(&th, &nm);
        if (!gs.ok()) return gs;
        HIP_TRY(guess_launch_features(stream_, na, d_atom_z_, d_x_, d_atom_inst_, d_atom0_, d_gfeat_));
        HIP_TRY(guess_launch_forward(stream_, na, th, nm, d_gfeat_, d_gout_));
        HIP_TRY(guess_launch_apply(stream_, ninst_, d_guess_usable_, d_atom0_, d_atom_sh0_, d_charge_, d_gout_, d_q0_));
        q0_dirty_ = true;
    }
    HIP_TRY(hipMemcpyAsync(d_qin_, d_q0_, nshells_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    HIP_TRY(hipEventRecord(ev_start_, stream_));
    if (serial_) HIP_TRY(hipDeviceSynchronize());
    for (auto& bk : buckets_) {
        HIP_TRY(hipStreamWaitEvent(bk.stream, ev_start_, 0));
        if (bk.in_setup) {
            HIP_TRY(hipEventRecord(bk.done, bk.stream));
            bk.inflight = true;
        }
    }
    return Status{};
}

// S and H0 of the bucket's instances on its stream: the molecular kernels, and
// the periodic (image-summed) ones when the bucket holds a periodic instance.
Status Batch::launch_integrals(Bucket& bk) {
    const DeviceTables t = dm_->tables();
    hipLaunchKernelGGL((k_overlap_h0_t<false, false>), dim3(bk.natoms), dim3(64), 0, bk.stream, db_, t,
                       (const int*)bk.atoms, (const double*)d_cn_, opt_.integral_cutoff);
    LAUNCH_CHECK();
    if (has_d_) {
        hipLaunchKernelGGL((k_overlap_h0_t<true, false>), dim3(bk.natoms), dim3(64), 0, bk.stream, db_, t,
                           (const int*)bk.atoms, (const double*)d_cn_, opt_.integral_cutoff);
        LAUNCH_CHECK();
    }
    if (bk.has_pbc) {
        hipLaunchKernelGGL((k_overlap_h0_t<false, true>), dim3(bk.natoms), dim3(64), 0, bk.stream, db_, t,
                           (const int*)bk.atoms, (const double*)d_cn_, kPbcAoCutoff);
        LAUNCH_CHECK();
        if (has_d_) {
            hipLaunchKernelGGL((k_overlap_h0_t<true, true>), dim3(bk.natoms), dim3(64), 0, bk.stream, db_, t,
                               (const int*)bk.atoms, (const double*)d_cn_, kPbcAoCutoff);
            LAUNCH_CHECK();
        }
    }
    return Status{};
}

// Setup of one bucket of at most kCholMaxDim AO on its stream: S and H0 over
// its atoms, S = L L^T and Cb = L^-T, failure flags (device state and a host
// copy), A0 = Cb^T H0 Cb.  Its `done` event is recorded by the caller once the
// start charges are ready.
Status Batch::launch_bucket_setup(Bucket& bk) {
    const DeviceTables t = dm_->tables();
    const int count = (int)bk.sel.size();
    const double one = 1.0, zero = 0.0;
    HIP_TRY(hipStreamWaitEvent(bk.stream, ev_cn_, 0));
    Status st = upload_active(bk, bk.sel, 0);
    if (!st.ok()) return st;
    if (bk.natoms > 0) {
        st = launch_integrals(bk);
        if (!st.ok()) return st;
    }
    hipLaunchKernelGGL(k_pad_diagonal, dim3(count), dim3(64), 0, bk.stream, db_, (const int*)bk.act);
    LAUNCH_CHECK();
    HIP_TRY(launch_cholesky_cb(bk.stream, count, (const int*)bk.act, (const int*)d_nao_, bk.ld,
                               bk.ptrs + (size_t)A_S * count, bk.ptrs + (size_t)A_L * count,
                               bk.ptrs + (size_t)A_CB * count, opt_.eigen_tolerance, bk.info));
    hipLaunchKernelGGL(k_mark_failed, dim3((count + 63) / 64), dim3(64), 0, bk.stream, (const int*)bk.act, count,
                       (const int*)bk.info, d_state_);
    LAUNCH_CHECK();
    {
        double* const* Cb = bk.ptrs + (size_t)A_CB * count;
        double* const* H0 = bk.ptrs + (size_t)A_H0 * count;
        double* const* T = bk.ptrs + (size_t)A_T * count;
        double* const* A0 = bk.ptrs + (size_t)A_A0 * count;
        const int n = bk.ld;
        RB_TRY(rocblas_dgemm_batched(bk.handle, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                                     (const double* const*)H0, n, (const double* const*)Cb, n, &zero,
                                     (double* const*)T, n, count));
        RB_TRY(rocblas_dgemm_batched(bk.handle, rocblas_operation_transpose, rocblas_operation_none, n, n, n, &one,
                                     (const double* const*)Cb, n, (const double* const*)T, n, &zero,
                                     (double* const*)A0, n, count));
    }
    hipLaunchKernelGGL(k_copy_words, dim3((count + 255) / 256), dim3(256), 0, bk.stream, (const unsigned int*)bk.info,
                       (unsigned int*)bk.h_aux, count);
    LAUNCH_CHECK();
    bk.in_setup = true;
    return Status{};
}

// The SCC runs each size bucket at its own pace: a bucket's next iteration is
// launched as soon as its previous one has finished, and the buckets never
// wait for each other.  Each iteration ends by gathering the states of the
// bucket's active instances into pinned host memory, behind an event; the host
// polls the events of the buckets in flight, reads the states of whichever
// bucket has finished, and launches its next iteration, or, once none of its
// instances iterates any more, its energy-weighted density and pair gradient.
Status Batch::scc_phase() {
    for (auto& bk : buckets_) {
        bk.active.clear();
        bk.iter = 0;
        bk.graph = false;
        bk.g_leave = false;
        bk.g_issued = bk.g_done = 0;
        if (bk.in_setup) continue;  // its setup is still in flight
        for (int i : bk.inst)
            if (h_state_[i].status == 0) bk.active.push_back(i);
        bk.inflight = false;
    }
    progress_ = std::getenv("GFN1_PROGRESS") != nullptr;
    t_scc_ = std::chrono::steady_clock::now();
    int setups = 0;
    for (auto& bk : buckets_) setups += bk.in_setup ? 1 : 0;
    if (!setups) setup_end_ = std::chrono::steady_clock::now();
    // Next work of a bucket whose stream is idle: replays while every
    // instance's iteration is plain, the host loop otherwise, the gradient once
    // nothing iterates any more.
    auto advance = [&](Bucket& bk) -> Status {
        if (bk.active.empty()) return launch_bucket_gradient(bk);
        bool plain = true;
        for (int i : bk.active) plain = plain && plain_iteration(h_state_[i], bk.kcap, bk.ld, broyden_memory_);
        if (plain && !bk.g_leave && bk.iter == 0 && graph_bucket(bk)) {
            Status st = capture_iteration(bk);
            if (!st.ok()) return st;
            bk.graph = true;
            bk.inflight = true;
            for (int r = 0; r < 2 && st.ok(); ++r) st = issue_replay(bk);
            return st;
        }
        bk.graph = false;
        // the iteration checks of finish_iteration that precede the next launch
        bool hot = false;
        for (int i : bk.active) hot = hot || h_state_[i].temperature > kWideSmearing;
        if (hot && bk.kcap < bk.ld) {
            Status gs = grow_kcap(bk, bk.ld);
            if (!gs.ok()) return gs;
        }
        ++launch_stats_.host_iterations;
        return launch_iteration(bk);
    };
    // largest matrices first: their chains are the longest, the small buckets fill in
    for (auto bit = buckets_.rbegin(); bit != buckets_.rend(); ++bit) {
        if (bit->in_setup) continue;
        Status st = advance(*bit);
        if (!st.ok()) return st;
        if (serial_) HIP_TRY(hipDeviceSynchronize());
    }
    while (true) {
        bool any = false, progressed = false;
        for (auto bit = buckets_.rbegin(); bit != buckets_.rend(); ++bit) {
            Bucket& bk = *bit;
            if (!bk.inflight) continue;
            any = true;
            if (bk.graph) {
                // read the states of every finished replay; keep two in flight
                while (bk.g_done < bk.g_issued) {
                    const hipError_t q = hipEventQuery(bk.g_ev[bk.g_done % 2]);
                    if (q == hipErrorNotReady) break;
                    HIP_TRY(q);
                    progressed = true;
                    bool to_host = false;
                    Status st = finish_replay(bk, to_host);
                    if (!st.ok()) return st;
                    if (!bk.g_leave && bk.g_issued - bk.g_done < 2) {
                        st = issue_replay(bk);
                        if (!st.ok()) return st;
                    }
                }
                if (bk.g_done < bk.g_issued) continue;
                // no replay in flight: the instances still iterating (if any)
                // need the host loop
                bk.inflight = false;
                bk.active.clear();
                for (int i : bk.inst)
                    if (h_sel_[i] && h_state_[i].status == 0) bk.active.push_back(i);
                bk.g_leave = true;
                Status st = advance(bk);
                if (!st.ok()) return st;
                if (serial_) HIP_TRY(hipDeviceSynchronize());
                continue;
            }
            const hipError_t q = hipEventQuery(bk.done);
            if (q == hipErrorNotReady) continue;
            HIP_TRY(q);
            bk.inflight = false;
            progressed = true;
            Status st;
            if (bk.in_setup) {
                // overlap matrices that are not positive definite (flags copied by the setup)
                bk.in_setup = false;
                const int* flag = (const int*)bk.h_aux;
                for (size_t k = 0; k < bk.sel.size(); ++k) {
                    const int i = bk.sel[k];
                    if (flag[k] != 0 && h_state_[i].status == 0) {
                        h_state_[i].status = 2;
                        if (!h_err_[i]) {
                            h_err_[i] = kErrInvalidInput;
                            h_errmsg_[i] = "overlap matrix is not positive definite";
                        }
                    }
                    if (h_state_[i].status == 0) bk.active.push_back(i);
                }
                if (--setups == 0) setup_end_ = std::chrono::steady_clock::now();
            } else {
                ++launch_stats_.state_reads;
                st = finish_iteration(bk);
                if (!st.ok()) return st;
            }
            st = advance(bk);
            if (!st.ok()) return st;
            if (serial_) HIP_TRY(hipDeviceSynchronize());
        }
        if (!any) break;
        if (!progressed) std::this_thread::yield();
    }
    for (int i = 0; i < ninst_; ++i)
        if (h_state_[i].status == 2 && !h_err_[i]) {
            h_err_[i] = kErrSccNotConverged;
            h_errmsg_[i] = "SCC did not converge";
        }
    return Status{};
}

// One SCC iteration of the bucket's active instances, ending with the copy of
// their states to the host and the bucket's `done` event.
Status Batch::launch_iteration(Bucket& bk) {
    const DeviceTables t = dm_->tables();
    const double one = 1.0, zero = 0.0;
    const int cnt = (int)bk.active.size();
    // kernels and library calls this iteration puts on the stream
    long long nk = 2;  // the two copies of upload_active
    Status st = upload_active(bk, bk.active, 0);
    if (!st.ok()) return st;
    int max_nsh = 0;
    for (int i : bk.active) max_nsh = std::max(max_nsh, h_nsh_[i]);
    const int n = bk.ld;
    // ---- Fock matrix from the input charges
    ++nk;
    hipLaunchKernelGGL(k_shell_potential, dim3((max_nsh + 127) / 128, cnt), dim3(128), 0, bk.stream, db_, t,
                       (const int*)bk.act, (const double*)d_qin_, d_vin_g_, d_vin_t_);
    LAUNCH_CHECK();
    // level shift F += b (S - 1/2 S P_prev S) for instances that carry a previous density
    std::vector<int> ls;
    for (int i : bk.active)
        if (h_state_[i].have_prev_p && h_state_[i].level_shift != 0.0) ls.push_back(i);
    // F itself is formed only for the result matrices and the level-shift rungs;
    // otherwise A comes from the reduced congruence below
    if (opt_.keep_matrices || !ls.empty()) {
        ++nk;
        hipLaunchKernelGGL(k_fock, dim3(n, cnt), dim3(256), 0, bk.stream, db_, (const int*)bk.act,
                           (const double*)d_vin_t_);
        LAUNCH_CHECK();
    }
    if (!ls.empty()) {
        // the staging buffers still feed the copy of the full active list
        HIP_TRY(hipStreamSynchronize(bk.stream));
        nk += 2;
        st = upload_active(bk, ls, 0);
        if (!st.ok()) return st;
        const int c = (int)ls.size();
        double* const* P = bk.ptrs + (size_t)A_P * c;
        double* const* S = bk.ptrs + (size_t)A_S * c;
        double* const* T = bk.ptrs + (size_t)A_T * c;
        double* const* Ab = bk.ptrs + (size_t)A_A * c;
        ++nk;
        RB_TRY(rocblas_dgemm_batched(bk.handle, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                                     (const double* const*)S, n, (const double* const*)P, n, &zero,
                                     (double* const*)T, n, c));
        ++nk;
        RB_TRY(rocblas_dgemm_batched(bk.handle, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                                     (const double* const*)T, n, (const double* const*)S, n, &zero,
                                     (double* const*)Ab, n, c));
        for (int k = 0; k < c; ++k) bk.h_aux[k] = h_state_[ls[k]].level_shift;
        ++nk;
        hipLaunchKernelGGL(k_copy_words, dim3((2 * c + 255) / 256), dim3(256), 0, bk.stream,
                           (const unsigned int*)bk.h_aux, (unsigned int*)bk.eig_e, 2 * c);
        LAUNCH_CHECK();
        ++nk;
        hipLaunchKernelGGL(k_level_shift, dim3(n, c), dim3(256), 0, bk.stream, db_, (const int*)bk.act,
                           (const double*)bk.eig_e, (double* const*)Ab);
        LAUNCH_CHECK();
        HIP_TRY(hipStreamSynchronize(bk.stream));
        nk += 2;
        st = upload_active(bk, bk.active, 0);
        if (!st.ok()) return st;
    }
    double* const* F = bk.ptrs + (size_t)A_F * cnt;
    double* const* Cb = bk.ptrs + (size_t)A_CB * cnt;  // L^-T (S = L L^T)
    double* const* Z = bk.ptrs + (size_t)A_C * cnt;    // eigenvectors of the reduced problem
    double* const* T = bk.ptrs + (size_t)A_T * cnt;
    double* const* A = bk.ptrs + (size_t)A_A * cnt;
    // ---- A = Cb^T F Cb (orthonormal basis; lower triangle and diagonal are used)
    if (!ls.empty()) {
        ++nk;
        RB_TRY(rocblas_dgemm_batched(bk.handle, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                                     (const double* const*)F, n, (const double* const*)Cb, n, &zero,
                                     (double* const*)T, n, cnt));
        ++nk;
        RB_TRY(rocblas_dgemm_batched(bk.handle, rocblas_operation_transpose, rocblas_operation_none, n, n, n, &one,
                                     (const double* const*)Cb, n, (const double* const*)T, n, &zero,
                                     (double* const*)A, n, cnt));
    } else {
        const int nt64 = (n + 63) / 64;
        ++nk;
        hipLaunchKernelGGL(k_reduce_fock, dim3(nt64, nt64, cnt), dim3(256), 0, bk.stream, db_, (const int*)bk.act,
                           (const double*)d_vin_t_, Cb, bk.ptrs + (size_t)A_L * cnt, bk.ptrs + (size_t)A_A0 * cnt,
                           A);
        LAUNCH_CHECK();
    }
    const int kc = bk.kcap;
    // matrices above kEigMaxDim go to rocSOLVER
    const bool own_solver = own_eigensolver(bk);
    if (own_solver) {
        // tridiagonal reduction, eigenvalues, occupations, occupied eigenvectors
        // (returned transposed in Z, kc columns)
        HIP_TRY(hipEventRecord(bk.ev0, bk.stream));
        ++nk;
        HIP_TRY(launch_tridiag(bk.stream, cnt, (const int*)bk.act, (const int*)d_nao_, n, A, bk.eig, bk.eig_e,
                               bk.tau));
        HIP_TRY(hipEventRecord(bk.ev1, bk.stream));
        bk.ev_pending = true;
        bk.ev_bytes = 0.0;
        for (int i : bk.active) {
            const double na = h_nao_[i];
            bk.ev_bytes += 8.0 * (na - 1.0) * na * (na + 1.0) / 6.0;
        }
        ++nk;
        HIP_TRY(launch_tridiag_eig(bk.stream, cnt, (const int*)bk.act, (const int*)d_nao_, n, (const double*)bk.eig,
                                   (const double*)bk.eig_e, d_eps_, (const int64_t*)d_vec0_, d_occ_, d_state_, Z,
                                   bk.lu, kc, opt_.keep_matrices != 0));
        nk += 2;
        HIP_TRY(launch_backtransform(bk.stream, cnt, (const int*)bk.act, (const int*)d_nao_, n, A,
                                     (const double*)bk.tau, Z, (const SccState*)d_state_, kc, bk.lu, bk.bt_cols));
        // C_k = Cb Z (into T); Z is stored transposed, element (m, j) at m * kc + j
        ++nk;
        HIP_TRY(launch_orbitals(bk.stream, cnt, (const int*)bk.act, (const int*)d_nao_, (const SccState*)d_state_, n,
                                kc, Cb, Z, kc, 1, T));
    } else {
        // rocSOLVER runs alone on the device, as in the setup phase
        HIP_TRY(hipDeviceSynchronize());
        const bool own_red = own_reduction(bk);
        Status st2 = own_red ? large_eigensolve(bk) : eigensolve(bk, cnt);
        if (!st2.ok()) return st2;
        ++nk;  // the rocSOLVER eigensolve counts once
        // rocSOLVER's convergence flags, read by the host once the iteration is done
        ++nk;
        hipLaunchKernelGGL(k_copy_words, dim3((cnt + 255) / 256), dim3(256), 0, bk.stream, (const unsigned int*)bk.info,
                           (unsigned int*)bk.h_aux, cnt);
        LAUNCH_CHECK();
        HIP_TRY(hipStreamSynchronize(bk.stream));
        ++nk;
        hipLaunchKernelGGL(k_fermi, dim3(cnt), dim3(256), 0, bk.stream, db_, (const int*)bk.act,
                           (const double*)bk.eig, n, d_eps_, d_occ_, d_state_);
        LAUNCH_CHECK();
        // the eigenvectors are column-major (element (m, j) at j * n + m): in place
        // in A from rocSOLVER's solvers, in the C slot from large_eigensolve
        ++nk;
        HIP_TRY(launch_orbitals(bk.stream, cnt, (const int*)bk.act, (const int*)d_nao_, (const SccState*)d_state_, n,
                                kc, Cb, own_red ? Z : A, 1, n, T));
    }
    // ---- P = C_k diag(f) C_k^T with the Mulliken row sums of S o P and H0 o P
    // (FP64 in every precision mode: the SCC energy and charges need it)
    double* const* P = bk.ptrs + (size_t)A_P * cnt;
    ++nk;
    HIP_TRY(launch_density(bk.stream, cnt, (const int*)bk.act, (const int*)d_nao_, (const SccState*)d_state_, n, kc,
                           T, (const double*)d_eps_, (const double*)d_occ_, (const int64_t*)d_vec0_, false, P,
                           bk.ptrs + (size_t)A_S * cnt, bk.ptrs + (size_t)A_H0 * cnt, (const int*)d_ao0_, d_rowS_,
                           d_rowH_, naos_));
    // ---- Mulliken charges, energy, convergence and mixing (above 512 AO the
    // row-sum blocks are first added up AO-parallel)
    int rs_blocks = 0;
    if (n > kEigMaxDim) {
        ++nk;
        hipLaunchKernelGGL(k_collapse_rowsums, dim3((n + 255) / 256, cnt), dim3(256), 0, bk.stream, db_,
                           (const int*)bk.act, d_rowS_, d_rowH_, naos_);
        LAUNCH_CHECK();
        rs_blocks = 1;
    }
    ++nk;
    hipLaunchKernelGGL(k_qout, dim3(cnt), dim3(128), 0, bk.stream, db_, t, (const int*)bk.act,
                       (const double*)d_rowS_, (const double*)d_rowH_, naos_, rs_blocks, d_qout_, d_state_);
    LAUNCH_CHECK();
    ++nk;
    hipLaunchKernelGGL(k_shell_potential, dim3((max_nsh + 127) / 128, cnt), dim3(128), 0, bk.stream, db_, t,
                       (const int*)bk.act, (const double*)d_qout_, d_vout_g_, d_vout_t_);
    LAUNCH_CHECK();
    // Broyden history longer than 32 needs global scratch
    int mmax = 0;
    for (int i : bk.active) mmax = std::max(mmax, std::min(h_state_[i].broyden_iter, broyden_memory_));
    const int beta_ld = std::max(mmax, 1);
    if (mmax > kBroydenLdsHistory) {
        size_t need = (size_t)cnt * (beta_ld + 1) * beta_ld;
        if (need > bk.beta_cap) {
            (void)hipFree(bk.beta);
            HIP_TRY(hipMalloc(&bk.beta, need * 2 * sizeof(double)));
            bk.beta_cap = need * 2;
        }
    }
    ++nk;
    hipLaunchKernelGGL(k_scc_update, dim3(cnt), dim3(256), 0, bk.stream, db_, t, (const int*)bk.act, prm_,
                       (const double*)d_qout_, (const double*)d_vout_g_, d_qin_, (const double*)d_q0_, d_state_,
                       d_brdf_, d_bru_, d_brom_, d_brql_, d_brdql_, (const int64_t*)d_broff_, bk.beta, beta_ld);
    LAUNCH_CHECK();
    ++nk;
    hipLaunchKernelGGL(k_gather_state, dim3((cnt + 63) / 64), dim3(64), 0, bk.stream, (const int*)bk.act, cnt,
                       (const SccState*)d_state_, bk.st_host);
    LAUNCH_CHECK();
    HIP_TRY(hipEventRecord(bk.done, bk.stream));
    launch_stats_.kernels += nk;
    launch_stats_.enqueues += nk;
    bk.inflight = true;
    ++bk.iter;
    return Status{};
}

// ---- Captured iterations
//
// A bucket solved by the own eigensolver iterates by replays of one captured
// iteration while its instances' iterations are plain ones (plain_iteration):
// the replay's kernels cover every position of the bucket, the gate kernel at
// its head marks the positions that do not iterate with -1, and the kernels
// skip those.  The host keeps two replays in flight per bucket and reads the
// states the first one gathered while the second runs, so the bucket's stream
// never waits for the host.  Each instance's arithmetic is the same as in a
// host-launched iteration: the gate only decides whether it runs.  Once an
// iteration that is not plain comes up (level shift, wide smearing beyond the
// eigenvector capacity, a long Broyden history), the bucket finishes the run
// in the host loop.
bool Batch::graph_bucket(const Bucket& bk) const {
    return opt_.scc_graphs && !serial_ && own_eigensolver(bk) && !bk.inst.empty();
}

// The kernels of one plain iteration over every position of the bucket, in
// the order of launch_iteration's own-eigensolver branch, states gathered into
// pinned slot `slot`.  Enqueued while the bucket's stream is being captured.
Status Batch::enqueue_plain_iteration(Bucket& bk, int slot) {
    const DeviceTables t = dm_->tables();
    const int cnt = (int)bk.inst.size();
    const int n = bk.ld, kc = bk.kcap;
    int max_nsh = 0;
    for (int i : bk.inst) max_nsh = std::max(max_nsh, h_nsh_[i]);
    auto table = [&](int role) { return (double* const*)(bk.g_ptrs + (size_t)role * cnt); };
    const int* act = bk.g_act;
    hipLaunchKernelGGL(k_gate_active, dim3((cnt + 255) / 256), dim3(256), 0, bk.stream, (const int*)bk.g_inst, cnt,
                       (const SccState*)d_state_, kc, n, broyden_memory_, bk.g_act);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_shell_potential, dim3((max_nsh + 127) / 128, cnt), dim3(128), 0, bk.stream, db_, t, act,
                       (const double*)d_qin_, d_vin_g_, d_vin_t_);
    LAUNCH_CHECK();
    if (opt_.keep_matrices) {
        hipLaunchKernelGGL(k_fock, dim3(n, cnt), dim3(256), 0, bk.stream, db_, act, (const double*)d_vin_t_);
        LAUNCH_CHECK();
    }
    const int nt64 = (n + 63) / 64;
    hipLaunchKernelGGL(k_reduce_fock, dim3(nt64, nt64, cnt), dim3(256), 0, bk.stream, db_, act,
                       (const double*)d_vin_t_, table(A_CB), table(A_L), table(A_A0), table(A_A));
    LAUNCH_CHECK();
    HIP_TRY(launch_tridiag(bk.stream, cnt, act, (const int*)d_nao_, n, table(A_A), bk.eig, bk.eig_e, bk.tau));
    HIP_TRY(launch_tridiag_eig(bk.stream, cnt, act, (const int*)d_nao_, n, (const double*)bk.eig,
                               (const double*)bk.eig_e, d_eps_, (const int64_t*)d_vec0_, d_occ_, d_state_, table(A_C),
                               bk.lu, kc, opt_.keep_matrices != 0));
    HIP_TRY(launch_backtransform(bk.stream, cnt, act, (const int*)d_nao_, n, table(A_A), (const double*)bk.tau,
                                 table(A_C), (const SccState*)d_state_, kc, bk.lu, bk.bt_cols));
    HIP_TRY(launch_orbitals(bk.stream, cnt, act, (const int*)d_nao_, (const SccState*)d_state_, n, kc, table(A_CB),
                            table(A_C), kc, 1, table(A_T)));
    HIP_TRY(launch_density(bk.stream, cnt, act, (const int*)d_nao_, (const SccState*)d_state_, n, kc, table(A_T),
                           (const double*)d_eps_, (const double*)d_occ_, (const int64_t*)d_vec0_, false, table(A_P),
                           table(A_S), table(A_H0), (const int*)d_ao0_, d_rowS_, d_rowH_, naos_));
    hipLaunchKernelGGL(k_qout, dim3(cnt), dim3(128), 0, bk.stream, db_, t, act, (const double*)d_rowS_,
                       (const double*)d_rowH_, naos_, 0, d_qout_, d_state_);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_shell_potential, dim3((max_nsh + 127) / 128, cnt), dim3(128), 0, bk.stream, db_, t, act,
                       (const double*)d_qout_, d_vout_g_, d_vout_t_);
    LAUNCH_CHECK();
    // the gate admits Broyden histories that fit in LDS only: no global scratch
    hipLaunchKernelGGL(k_scc_update, dim3(cnt), dim3(256), 0, bk.stream, db_, t, act, prm_, (const double*)d_qout_,
                       (const double*)d_vout_g_, d_qin_, (const double*)d_q0_, d_state_, d_brdf_, d_bru_, d_brom_,
                       d_brql_, d_brdql_, (const int64_t*)d_broff_, bk.beta, 1);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gather_state, dim3((cnt + 63) / 64), dim3(64), 0, bk.stream, (const int*)bk.g_inst, cnt,
                       (const SccState*)d_state_, bk.g_st[slot]);
    LAUNCH_CHECK();
    return Status{};
}

// Pointer tables of every position (parity 0) and the graphs of both state
// slots.  Kept across runs; captured again when the eigenvector capacity or
// its scratch changed.  The bucket's stream is idle.
Status Batch::capture_iteration(Bucket& bk) {
    if (bk.gexec[0] && bk.gexec[1] && bk.g_kcap == bk.kcap && bk.g_lu == bk.lu && bk.g_bt_cols == bk.bt_cols)
        return Status{};
    const auto t0 = std::chrono::steady_clock::now();
    for (int slot = 0; slot < 2; ++slot)
        if (bk.gexec[slot]) {
            HIP_TRY(hipGraphExecDestroy(bk.gexec[slot]));
            bk.gexec[slot] = nullptr;
        }
    const size_t mat = (size_t)bk.ld * bk.ld;
    const size_t count = bk.inst.size();
    double** hp = bk.h_ptrs;
    for (size_t k = 0; k < count; ++k) {
        auto at = [&](int role) { return bk.arena + ((size_t)role * count + k) * mat; };
        hp[(size_t)A_F * count + k] = at(R_F);
        hp[(size_t)A_CB * count + k] = at(R_CB0);
        hp[(size_t)A_C * count + k] = at(R_CB1);
        hp[(size_t)A_T * count + k] = at(R_T);
        hp[(size_t)A_A * count + k] = at(R_A);
        hp[(size_t)A_P * count + k] = at(R_P);
        hp[(size_t)A_S * count + k] = at(R_S);
        hp[(size_t)A_X * count + k] = at(R_W);
        hp[(size_t)A_C32 * count + k] = (double*)((float*)at(R_A) + (size_t)bk.ld * bk.kcap);
        hp[(size_t)A_H0 * count + k] = at(R_H0);
        hp[(size_t)A_L * count + k] = at(R_L);
        hp[(size_t)A_A0 * count + k] = at(R_A0);
        bk.h_act[k] = bk.inst[k];
    }
    const int wp = (int)(kActRoles * count * sizeof(double*) / sizeof(unsigned int));
    hipLaunchKernelGGL(k_copy_words, dim3((wp + 255) / 256), dim3(256), 0, bk.stream, (const unsigned int*)hp,
                       (unsigned int*)bk.g_ptrs, wp);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_copy_words, dim3(((int)count + 255) / 256), dim3(256), 0, bk.stream,
                       (const unsigned int*)bk.h_act, (unsigned int*)bk.g_inst, (int)count);
    LAUNCH_CHECK();
    // the staging buffers are free again once the copies have run
    HIP_TRY(hipStreamSynchronize(bk.stream));
    for (int slot = 0; slot < 2; ++slot) {
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(bk.stream, hipStreamCaptureModeRelaxed));
        Status st = enqueue_plain_iteration(bk, slot);
        const hipError_t end = hipStreamEndCapture(bk.stream, &graph);
        if (!st.ok()) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        HIP_TRY(end);
        size_t nodes = 0;
        hipError_t e = hipGraphGetNodes(graph, nullptr, &nodes);
        if (e == hipSuccess) e = hipGraphInstantiate(&bk.gexec[slot], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_TRY(e);
        bk.g_nodes = (int)nodes;
    }
    bk.g_kcap = bk.kcap;
    bk.g_lu = bk.lu;
    bk.g_bt_cols = bk.bt_cols;
    ++launch_stats_.graphs_captured;
    launch_stats_.capture_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return Status{};
}

Status Batch::issue_replay(Bucket& bk) {
    const int slot = (int)(bk.g_issued % 2);
    HIP_TRY(hipGraphLaunch(bk.gexec[slot], bk.stream));
    HIP_TRY(hipEventRecord(bk.g_ev[slot], bk.stream));
    ++bk.g_issued;
    ++launch_stats_.graph_replays;
    ++launch_stats_.enqueues;
    launch_stats_.kernels += bk.g_nodes;
    return Status{};
}

// States gathered by the oldest replay in flight (its event has completed):
// the same checks as after a host-launched iteration for the instances that
// iterated in it.  switch_to_host is set when an instance still iterating needs
// an iteration that is not plain.
Status Batch::finish_replay(Bucket& bk, bool& switch_to_host) {
    const SccState* g = bk.g_st[bk.g_done % 2];
    ++bk.g_done;
    ++bk.iter;
    ++launch_stats_.state_reads;
    int plain = 0;
    switch_to_host = false;
    for (size_t k = 0; k < bk.inst.size(); ++k) {
        const int i = bk.inst[k];
        if (!h_sel_[i]) continue;
        const bool iterated = g[k].total_iter != h_state_[i].total_iter;
        h_state_[i] = g[k];
        if (iterated) {
            if (h_state_[i].kneed_full > bk.kcap)
                return Status::error(kErrLinearAlgebra, "more partially occupied orbitals than the eigenvector capacity");
            if (h_state_[i].eig_fail)
                return Status::error(kErrLinearAlgebra, "eigenvalue cluster too large for the batched solver");
        }
        if (h_state_[i].status != 0) continue;
        if (plain_iteration(h_state_[i], bk.kcap, bk.ld, broyden_memory_)) ++plain;
        else switch_to_host = true;
    }
    if (progress_) {
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_scc_).count();
        std::fprintf(stderr, "scc replay %d of the %d-AO bucket: %d instances still iterating, %.1f ms\n", bk.iter,
                     bk.ld, plain, ms);
        std::fflush(stderr);
    }
    if (plain == 0) bk.g_leave = true;
    if (switch_to_host) bk.g_leave = true;
    return Status{};
}

// ---- Autotuning of the back-transformation width
//
// After a run, every own-eigensolver bucket times launch_backtransform at each
// admissible width (16, 32, and 64 eigenvectors per work-group up to 256 rows)
// on the reflectors and T factors its instances left, one untimed launch and
// three timed ones per width (HIP events on the bucket's stream), and keeps the
// fastest.  The eigenvector buffers it overwrites are dead between runs (every
// iteration recomputes them), and all widths give the same result bits, so the
// choice never changes a result.
Status Batch::autotune() {
    if (!ran_) return Status::error(kErrInvalidInput, "autotune times the kernels on a finished run's data");
    HIP_TRY(hipDeviceSynchronize());
    tuning_.clear();
    static const int widths[3] = {16, 32, 64};
    for (auto& bk : buckets_) {
        if (!own_eigensolver(bk)) continue;
        std::vector<int> list;
        for (int i : bk.sel)
            if (h_state_[i].total_iter > 0) list.push_back(i);
        if (list.empty()) continue;
        Status st = upload_active(bk, list, 0);
        if (!st.ok()) return st;
        const int cnt = (int)list.size(), n = bk.ld;
        double* const* A = bk.ptrs + (size_t)A_A * cnt;
        double* const* Z = bk.ptrs + (size_t)A_C * cnt;
        TuneRecord rec;
        rec.ld = n;
        rec.count = cnt;
        rec.cols = bk.bt_cols;
        double best = 0.0;
        for (int w = 0; w < 3; ++w) {
            if (!backtransform_width_ok(n, widths[w])) continue;
            HIP_TRY(launch_backtransform(bk.stream, cnt, (const int*)bk.act, (const int*)d_nao_, n, A,
                                         (const double*)bk.tau, Z, (const SccState*)d_state_, bk.kcap, bk.lu,
                                         widths[w]));
            HIP_TRY(hipEventRecord(bk.ev0, bk.stream));
            for (int r = 0; r < 3; ++r)
                HIP_TRY(launch_backtransform(bk.stream, cnt, (const int*)bk.act, (const int*)d_nao_, n, A,
                                             (const double*)bk.tau, Z, (const SccState*)d_state_, bk.kcap, bk.lu,
                                             widths[w]));
            HIP_TRY(hipEventRecord(bk.ev1, bk.stream));
            HIP_TRY(hipEventSynchronize(bk.ev1));
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, bk.ev0, bk.ev1));
            rec.ms[w] = ms / 3.0;
            if (best == 0.0 || rec.ms[w] < best) {
                best = rec.ms[w];
                rec.cols = widths[w];
            }
        }
        bk.bt_cols = rec.cols;
        tuning_.push_back(rec);
    }
    // the events of the host loop's timing must not pair with these
    for (auto& bk : buckets_) bk.ev_pending = false;
    return Status{};
}

// Host side of a finished iteration: states, kernel timing, checks, next active list.
Status Batch::finish_iteration(Bucket& bk) {
    for (size_t j = 0; j < bk.active.size(); ++j) h_state_[bk.active[j]] = bk.st_host[j];
    if (bk.ev_pending) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, bk.ev0, bk.ev1));
        tri_ms_ += ms;
        tri_bytes_ += bk.ev_bytes;
        ++tri_launches_;
        bk.ev_pending = false;
    }
    if (!own_eigensolver(bk) && !bk.active.empty()) {
        const int* info = (const int*)bk.h_aux;
        for (size_t j = 0; j < bk.active.size(); ++j)
            if (info[j] != 0) return Status::error(kErrLinearAlgebra, "eigensolver did not converge");
    }
    // the last ladder rung smears at >= 3000 K: keep every eigenvector from then on
    bool hot = false;
    for (int i : bk.active) hot = hot || h_state_[i].temperature > kWideSmearing;
    if (hot && bk.kcap < bk.ld) {
        Status gs = grow_kcap(bk, bk.ld);
        if (!gs.ok()) return gs;
    }
    std::vector<int> next;
    for (int i : bk.active) {
        if (h_state_[i].kneed_full > bk.kcap)
            return Status::error(kErrLinearAlgebra, "more partially occupied orbitals than the eigenvector capacity");
        if (h_state_[i].eig_fail)
            return Status::error(kErrLinearAlgebra, "eigenvalue cluster too large for the batched solver");
        if (h_state_[i].status == 0) next.push_back(i);
    }
    bk.active.swap(next);
    if (progress_) {
        size_t left = 0;
        for (auto& b : buckets_) left += b.active.size();
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_scc_).count();
        std::fprintf(stderr, "scc iteration %d of the %d-AO bucket: %zu instances still iterating, %.1f ms\n", bk.iter,
                     bk.ld, left, ms);
        std::fflush(stderr);
    }
    return Status{};
}

// Energy-weighted density W = C_k diag(eps f) C_k^T of the bucket's converged
// instances, from the orbitals of their last iteration (kept in T), and the
// pair part of the electronic gradient of the bucket's atoms, on the bucket's
// stream.  Runs while other buckets are still iterating.
Status Batch::launch_bucket_gradient(Bucket& bk) {
    const DeviceTables t = dm_->tables();
    if (opt_.compute_gradient && bk.npinst > 0) {
        // the Ewald gradient needs only the converged charges: it runs on its
        // own stream beside the density and pair-gradient kernels below
        HIP_TRY(hipEventRecord(ev_ew_q_, bk.stream));
        HIP_TRY(hipStreamWaitEvent(ew_stream_, ev_ew_q_, 0));
        HIP_TRY(launch_ewald_grad(ew_stream_, db_, t, bk.pinst, bk.npinst, bk.max_ng, bk.patoms, bk.npatoms,
                                  (const double*)d_qout_, d_gew_));
    }
    std::vector<int> done;
    for (int i : bk.inst)
        if (h_state_[i].status == 1) done.push_back(i);
    const int cnt = (int)done.size();
    if (cnt) {
        Status st = upload_active(bk, done, 0);
        if (!st.ok()) return st;
        const double one = 1.0, zero = 0.0;
        const int n = bk.ld, kc = bk.kcap;
        double* const* T = bk.ptrs + (size_t)A_T * cnt;
        double* const* A = bk.ptrs + (size_t)A_A * cnt;
        double* const* W = bk.ptrs + (size_t)A_X * cnt;
        const int nt = (n + 31) / 32;
        if (opt_.mixed) {
            double* const* C32 = bk.ptrs + (size_t)A_C32 * cnt;
            double* const* X32 = bk.ptrs + (size_t)A_C * cnt;  // the eigenvector buffer is free after the SCC
            hipLaunchKernelGGL(k_operands_fp32, dim3(kc, cnt), dim3(256), 0, bk.stream, db_, (const int*)bk.act, T,
                               A, C32, (const double*)d_eps_, (const double*)d_occ_, kc, 1);
            LAUNCH_CHECK();
            const float onef = 1.0f, zerof = 0.0f;
            RB_TRY(rocblas_sgemm_batched(bk.handle, rocblas_operation_none, rocblas_operation_transpose, n, n, kc,
                                         &onef, (const float* const*)A, n, (const float* const*)C32, n, &zerof,
                                         (float* const*)X32, n, cnt));
            hipLaunchKernelGGL(k_symmetrize_fp32, dim3(nt, nt, cnt), dim3(256), 0, bk.stream, db_,
                               (const int*)bk.act, X32, W);
            LAUNCH_CHECK();
        } else {
            HIP_TRY(launch_density(bk.stream, cnt, (const int*)bk.act, (const int*)d_nao_, (const SccState*)d_state_,
                                   n, kc, T, (const double*)d_eps_, (const double*)d_occ_, (const int64_t*)d_vec0_,
                                   true, W, nullptr, nullptr, (const int*)d_ao0_, nullptr, nullptr, naos_));
        }
    }
    // the periodic gradient ignores the include flags and always has every
    // term (pbc/gradient.rs:75-163)
    const bool el_on = opt_.include_hamiltonian || opt_.include_scc || bk.has_pbc;
    if (opt_.compute_gradient && el_on && bk.natoms > 0) {
        const int cn_on = opt_.enable_cn_hamiltonian;
        auto pairs = opt_.mixed ? k_grad_pairs_t<false, true, false> : k_grad_pairs_t<false, false, false>;
        hipLaunchKernelGGL(pairs, dim3(bk.natoms), dim3(64), 0, bk.stream, db_, t, (const int*)bk.atoms,
                           (const double*)d_cn_, (const double*)d_vout_t_, opt_.include_scc, cn_on,
                           opt_.integral_cutoff, bk.gpair, bk.dedcn, (double*)nullptr);
        LAUNCH_CHECK();
        if (has_d_) {
            auto dpairs = opt_.mixed ? k_grad_pairs_t<true, true, false> : k_grad_pairs_t<true, false, false>;
            hipLaunchKernelGGL(dpairs, dim3(bk.natoms), dim3(64), 0, bk.stream, db_, t, (const int*)bk.atoms,
                               (const double*)d_cn_, (const double*)d_vout_t_, opt_.include_scc, cn_on,
                               opt_.integral_cutoff, bk.gpair, bk.dedcn, (double*)nullptr);
            LAUNCH_CHECK();
        }
        if (bk.has_pbc) {
            // symmetry-unique periodic pairs with their ket shares in d_ket_ (LDS
            // accumulator of 4 doubles per atom of the largest periodic instance)
            const size_t ket_lds = d_ket_ ? (size_t)4 * max_pnat_ * sizeof(double) : 0;
            auto ppairs = opt_.mixed ? k_grad_pairs_t<false, true, true> : k_grad_pairs_t<false, false, true>;
            hipLaunchKernelGGL(ppairs, dim3(bk.natoms), dim3(64), ket_lds, bk.stream, db_, t, (const int*)bk.atoms,
                               (const double*)d_cn_, (const double*)d_vout_t_, 1, cn_on, kPbcAoCutoff, bk.gpair,
                               bk.dedcn, d_ket_);
            LAUNCH_CHECK();
            if (has_d_) {
                auto pdpairs = opt_.mixed ? k_grad_pairs_t<true, true, true> : k_grad_pairs_t<true, false, true>;
                hipLaunchKernelGGL(pdpairs, dim3(bk.natoms), dim3(64), ket_lds, bk.stream, db_, t,
                                   (const int*)bk.atoms, (const double*)d_cn_, (const double*)d_vout_t_, 1, cn_on,
                                   kPbcAoCutoff, bk.gpair, bk.dedcn, d_ket_);
                LAUNCH_CHECK();
            }
        }
    }
    return Status{};
}

// After every bucket has finished: the per-bucket pair sums go to the per-atom
// arrays, then the atom-wise terms and the CN chain rule.
Status Batch::gradient_phase() {
    const DeviceTables t = dm_->tables();
    if (!opt_.compute_gradient || natoms_ == 0) return Status{};
    const int na = natoms_;
    HIP_TRY(hipMemsetAsync(d_gel_, 0, 3 * na * sizeof(double), stream_));
    const bool el_on = opt_.include_hamiltonian || opt_.include_scc || !h_pinst_.empty();
    const int cn_on = opt_.enable_cn_hamiltonian;
    if (el_on) {
        for (auto& bk : buckets_) {
            if (bk.natoms == 0) continue;
            hipLaunchKernelGGL(k_scatter_atoms, dim3((bk.natoms + 127) / 128), dim3(128), 0, stream_,
                               (const int*)bk.atoms, bk.natoms, (const double*)bk.gpair, (const double*)bk.dedcn,
                               d_gel_, d_dedcn_h_);
            LAUNCH_CHECK();
        }
        if (!h_patoms_.empty()) {
            // the ket shares of the periodic pair gradient and the Ewald gradient
            HIP_TRY(hipEventRecord(ev_ew_, ew_stream_));
            HIP_TRY(hipStreamWaitEvent(stream_, ev_ew_, 0));
            const int np = (int)h_patoms_.size();
            hipLaunchKernelGGL(k_periodic_merge, dim3((np + 127) / 128), dim3(128), 0, stream_, db_,
                               (const int*)d_patoms_, np, (const double*)d_ket_, (const double*)d_gew_, d_gel_,
                               d_dedcn_h_);
            LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_grad_atoms, dim3((na + 127) / 128), dim3(128), 0, stream_, db_, t,
                           (const double*)d_qout_, cn_on, opt_.include_scc, d_dedcn_h_, d_gel_, na);
        LAUNCH_CHECK();
        if (cn_on) {
            hipLaunchKernelGGL(k_cn_chain, dim3((na + 127) / 128), dim3(128), 0, stream_, db_, t,
                               (const double*)d_dedcn_h_, opt_.coordination_cutoff, 1e-12, d_gel_, na);
            LAUNCH_CHECK();
            if (!h_patoms_.empty()) {
                hipLaunchKernelGGL(k_cn_chain_pbc, dim3((int)h_patoms_.size()), dim3(64), 0, stream_, db_, t,
                                   (const int*)d_patoms_, (const double*)d_dedcn_h_, opt_.coordination_cutoff,
                                   1e-12, d_gel_);
                LAUNCH_CHECK();
            }
        }
    }
    HIP_TRY(hipStreamSynchronize(stream_));
    return Status{};
}

Status Batch::download(std::vector<InstanceResult>& out) {
    out.assign(ninst_, InstanceResult{});
    const int na = natoms_;
    std::vector<double> cn(na), gel(3 * na), grep(3 * na), gdisp(3 * na), ghal(3 * na), qout(nshells_),
        vout(nshells_), erep(ninst_), edisp(ninst_), ehal(ninst_), eps(h_vec0_[ninst_]), occ(h_vec0_[ninst_]);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(copy_bytes(h_state_.data(), d_state_, ninst_ * sizeof(SccState), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(cn.data(), d_cn_, na * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(gel.data(), d_gel_, 3 * na * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(grep.data(), d_grep_, 3 * na * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(gdisp.data(), d_gdisp_, 3 * na * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(ghal.data(), d_ghal_, 3 * na * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(qout.data(), d_qout_, nshells_ * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(vout.data(), d_vout_t_, nshells_ * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(erep.data(), d_erep_, ninst_ * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(edisp.data(), d_edisp_, ninst_ * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(ehal.data(), d_ehal_, ninst_ * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(eps.data(), d_eps_, eps.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(copy_bytes(occ.data(), d_occ_, occ.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double*> hmat((size_t)kRoles * ninst_);
    HIP_TRY(copy_bytes(hmat.data(), d_mat_, hmat.size() * sizeof(double*), hipMemcpyDeviceToHost));
    for (int i = 0; i < ninst_; ++i) {
        InstanceResult& r = out[i];
        const SccState& s = h_state_[i];
        const int a0 = h_atom0_[i], nat = h_atom0_[i + 1] - a0;
        const int s0 = h_sh0_[i], nsh = h_nsh_[i];
        const int n = h_nao_[i], ld = h_ld_[i];
        r.nat = nat;
        r.nsh = nsh;
        r.nao = n;
        r.iterations = s.iter;
        r.converged = s.status == 1;
        r.rms = s.rms;
        r.status = h_err_[i] ? h_err_[i] : (s.status == 1 ? kOk : kErrSccNotConverged);
        r.message = h_errmsg_[i];
        if (s.status == 3) {
            r.status = kErrInvalidInput;
            r.message = "instance not selected in the last run";
        }
        if (r.status != kOk) continue;
        const double e_disp = opt_.enable_dispersion ? edisp[i] : 0.0;
        const double total_internal = s.e_el + s.e2 + s.e3 + erep[i] + e_disp + ehal[i];
        for (double& e : r.energies) e = 0.0;
        r.energies[0] = total_internal + s.ts;
        r.energies[1] = total_internal;
        r.energies[2] = erep[i];
        r.energies[3] = s.e_el;
        r.energies[4] = s.e2;
        r.energies[5] = s.e3;
        r.energies[7] = e_disp;
        r.energies[8] = ehal[i];
        r.energies[14] = s.ts;
        r.fermi_level = s.mu;
        r.temperature = s.temperature;
        r.nelec = s.nelec;
        r.shell_charges.assign(qout.begin() + s0, qout.begin() + s0 + nsh);
        r.shell_potential.assign(vout.begin() + s0, vout.begin() + s0 + nsh);
        r.cn.assign(cn.begin() + a0, cn.begin() + a0 + nat);
        r.atomic_charges.assign(nat, 0.0);
        for (int k = 0; k < nsh; ++k) r.atomic_charges[h_sh_atom_[s0 + k] - a0] += r.shell_charges[k];
        for (int d = 0; d < 3; ++d) {
            double v = 0.0;
            for (int a = 0; a < nat; ++a) v += h_x_[3 * (a0 + a) + d] * r.atomic_charges[a];
            r.dipole[d] = v;
        }
        r.eps.assign(eps.begin() + h_vec0_[i], eps.begin() + h_vec0_[i] + n);
        r.occ.assign(occ.begin() + h_vec0_[i], occ.begin() + h_vec0_[i] + n);
        auto slice3 = [&](const std::vector<double>& g, bool on) {
            std::vector<double> v(3 * nat, 0.0);
            if (on)
                for (int k = 0; k < 3 * nat; ++k) v[k] = g[3 * a0 + k];
            return v;
        };
        if (opt_.compute_gradient) {
            // the periodic gradient has every term whatever the include flags
            // (pbc/gradient.rs:75-163, gradient.rs:74-120)
            const bool all = periodic(i);
            const bool el_on = opt_.include_hamiltonian || opt_.include_scc || all;
            r.g_el = slice3(gel, el_on);
            r.g_rep = slice3(grep, opt_.include_repulsion || all);
            r.g_disp = slice3(gdisp, (opt_.include_dispersion || all) && opt_.enable_dispersion);
            r.g_hal = slice3(ghal, opt_.include_halogen || all);
            r.gradient.assign(3 * nat, 0.0);
            for (int k = 0; k < 3 * nat; ++k) r.gradient[k] = r.g_el[k] + r.g_rep[k] + r.g_disp[k] + r.g_hal[k];
        }
        if (opt_.keep_matrices) {
            auto grab = [&](int role, std::vector<double>& dst) -> Status {
                std::vector<double> full((size_t)ld * ld);
                HIP_TRY(copy_bytes(full.data(), hmat[(size_t)role * ninst_ + i], full.size() * sizeof(double),
                                  hipMemcpyDeviceToHost));
                dst.assign((size_t)n * n, 0.0);
                for (int c = 0; c < n; ++c)
                    for (int rr = 0; rr < n; ++rr) dst[(size_t)rr * n + c] = full[(size_t)c * ld + rr];
                return Status{};
            };
            Status st;
            if (!(st = grab(R_S, r.S)).ok()) return st;
            if (!(st = grab(R_H0, r.H0)).ok()) return st;
            if (!(st = grab(R_F, r.F)).ok()) return st;
            if (!(st = grab(R_P, r.P)).ok()) return st;
            if (!(st = grab(R_W, r.W)).ok()) return st;
        }
    }
    return Status{};
}

}  // namespace gfn1
