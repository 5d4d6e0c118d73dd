// Kernel entry points of kernels.hip (launched by the batch driver).
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace gfn1 {

// SCC controls shared by every instance of a batch: the convergence thresholds
// and the robust-fallback ladder (electronic.rs:1413-1419).
struct SccParams {
    double energy_tolerance;
    double charge_tolerance;
    int32_t broyden_memory;   // BroydenMixer memory (history modulus)
    int32_t broyden_slots;    // allocated history slots per instance
    int32_t omega_stride;     // per-instance stride of the Broyden weights (broyden_slots padded)
    int32_t rung_broyden[4];
    double rung_mixing[4];
    double rung_level_shift[4];
    double rung_temperature[4];
    int32_t rung_max_iter[4];
};

// Molecular atoms only; the listed periodic atoms go to the _pbc kernels (one
// 64-lane work-group per atom, launched after their molecular counterparts).
__global__ void k_coordination(DeviceBatch b, DeviceTables t, double cn_cut, int cn_on, double* cn, double* cn_d3,
                               double* gw, double* dgw, int natoms);
__global__ void k_coordination_pbc(DeviceBatch b, DeviceTables t, const int* patoms, double cn_cut, int cn_on,
                                   double* cn, double* cn_d3, double* gw, double* dgw);
// S and H0 of the listed atoms' instances; PBC selects the periodic instances
// (the Gamma-point image sums), the other instantiation the molecular ones.
template <bool DCLASS, bool PBC>
__global__ void k_overlap_h0_t(DeviceBatch b, DeviceTables t, const int* atoms, const double* cn, double cutoff);
// Padding diagonal of the listed instances (act[blockIdx.x]).
__global__ void k_pad_diagonal(DeviceBatch b, const int* act);
// st[act[j]].status = 2 where flag[j] != 0 and the instance is still iterating.
__global__ void k_mark_failed(const int* act, int n, const int* flag, SccState* st);
__global__ void k_make_cb(DeviceBatch b, double* const* Linv, const int* act, double eig_tol, int* fail);
__global__ void k_shell_potential(DeviceBatch b, DeviceTables t, const int* act, const double* q, double* vg,
                                  double* vt);
__global__ void k_fock(DeviceBatch b, const int* act, const double* vt);
__global__ void k_level_shift(DeviceBatch b, const int* act, const double* shift, double* const* SPS);
__global__ void k_fermi(DeviceBatch b, const int* act, const double* eig, int eig_stride, double* eps_all,
                        double* occ_all, SccState* st);
__global__ void k_reduce_fock(DeviceBatch b, const int* act, const double* vt, double* const* Cbp, double* const* Lp,
                              double* const* A0p, double* const* Ap);
__global__ void k_operands_fp32(DeviceBatch b, const int* act, double* const* Cp, double* const* A32p,
                                double* const* C32p, const double* eps_all, const double* occ_all, int k,
                                int weighted);
__global__ void k_symmetrize_fp32(DeviceBatch b, const int* act, double* const* Xp, double* const* Mp);
__global__ void k_qout(DeviceBatch b, DeviceTables t, const int* act, const double* rowS, const double* rowH,
                       int stride, int blocks, double* qout, SccState* st);
__global__ void k_collapse_rowsums(DeviceBatch b, const int* act, double* rowS, double* rowH, int stride);
__global__ void k_scc_update(DeviceBatch b, DeviceTables t, const int* act, SccParams prm, const double* qout,
                             const double* vg_out, double* qin, const double* q0, SccState* st, double* br_df,
                             double* br_u, double* br_omega, double* br_qlast, double* br_dqlast,
                             const int64_t* br_off, double* beta_scratch, int beta_ld);
__global__ void k_geometry_terms(DeviceBatch b, DeviceTables t, const double* gw, const double* cn_d3, int disp_on,
                                 double* e_rep_atom, double* e_disp_atom, double* g_rep, double* g_disp,
                                 double* dedcn_disp, int natoms);
__global__ void k_geometry_terms_pbc(DeviceBatch b, DeviceTables t, const int* patoms, const double* gw, int disp_on,
                                     double* e_rep_atom, double* e_disp_atom, double* g_rep, double* g_disp,
                                     double* dedcn_disp, int natoms);
__global__ void k_cn_chain(DeviceBatch b, DeviceTables t, const double* dedcn, double cut, double reps, double* g,
                           int natoms);
__global__ void k_cn_chain_pbc(DeviceBatch b, DeviceTables t, const int* patoms, const double* dedcn, double cut,
                               double reps, double* g);
__global__ void k_halogen(DeviceBatch b, DeviceTables t, double* e_hal, double* g_hal);
__global__ void k_instance_sum(DeviceBatch b, const double* per_atom, double* out);
// Pair part of the electronic gradient for the listed atoms: g_pair[3 j + d] and
// dedcn_h[j] belong to atom atoms[j].  Periodic instantiation with ket_rows:
// each symmetry-unique (atom, atom, image) pair once, the ket atom's share in
// the instance's ket-share rows (launch with 4 nat doubles of dynamic shared
// memory, nat the largest periodic atom count), summed by k_periodic_merge;
// ket_rows = nullptr evaluates every pair from both sides.
template <bool DCLASS, bool MIXED, bool PBC>
__global__ void k_grad_pairs_t(DeviceBatch b, DeviceTables t, const int* atoms, const double* cn, const double* vsh,
                               int scc_on, int cn_on, double cutoff, double* g_pair, double* dedcn_h,
                               double* ket_rows);
__global__ void k_periodic_merge(DeviceBatch b, const int* patoms, int npatoms, const double* ket_rows,
                                 const double* g_ew, double* g, double* dedcn);
// Per-atom scatter of listed sums: g[3 atoms[j] + d] = gl[3 j + d], dc[atoms[j]] = dcl[j].
__global__ void k_scatter_atoms(const int* atoms, int n, const double* gl, const double* dcl, double* g, double* dc);
// out[j] = st[act[j]] for j < n (out may be coherent pinned host memory).
__global__ void k_gather_state(const int* act, int n, const SccState* st, SccState* out);
// dst[i] = src[i] for i < n 32-bit words (src may be coherent pinned host memory).
__global__ void k_gate_active(const int* inst, int count, const SccState* st, int kcap, int ld, int broyden_memory,
                              int* act);
__global__ void k_copy_words(const unsigned int* src, unsigned int* dst, int n);
__global__ void k_grad_atoms(DeviceBatch b, DeviceTables t, const double* q, int cn_on, int scc_on,
                             double* dedcn_h, double* g_el, int natoms);

}  // namespace gfn1
