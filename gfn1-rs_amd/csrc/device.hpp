// Device-side data layout shared by the kernels and the batch driver.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "model.hpp"

namespace gfn1 {

// Pointers fetched from device-side pointer arrays are generic to the compiler,
// which then emits FLAT accesses that also count against the LDS wait counter.
// Casting them to the global address space gives plain global loads/stores.
#define GFN1_GLOBAL __attribute__((address_space(1)))
template <class T>
using gptr = GFN1_GLOBAL T*;

// Padded AO dimension of an instance's dense matrices; padded rows/columns carry
// an identity overlap and a diagonal Fock value of kPadEnergy, so they decouple
// exactly and their eigenvalues sort above every physical orbital.
constexpr double kPadEnergy = 100.0;
constexpr int kPadMultiple = 32;

// Per-instance SCC control state (one entry per instance, device resident).
// One instance's SCC state, padded to 256 bytes: the size buckets run on
// concurrent streams, and the XCDs' L2 caches are only made coherent at kernel
// boundaries, so no two instances may share a cache line.
struct alignas(256) SccState {
    int32_t status;        // 0 running, 1 converged, 2 failed, 3 not selected for the run
    int32_t rung;          // robust-fallback ladder rung (electronic.rs:1413-1419)
    int32_t iter;          // iterations done on the current rung
    int32_t total_iter;
    int32_t have_energy;   // a previous iteration energy exists on this rung
    int32_t have_prev_p;   // level-shift projector available
    int32_t broyden_iter;  // BroydenMixer::iter
    int32_t kneed;         // orbitals with non-negligible occupation
    int32_t eig_fail;
    int32_t c_parity;      // which coefficient buffer holds the current orbitals
    int32_t kneed_full;    // kneed before clamping to the eigenvector capacity
    int32_t pad1;
    double mixing, level_shift, temperature, kt;
    double e_last, energy, rms, ts, mu, e_el, e2, e3;
    double nelec;
};

// Smearing temperature (K) from which every eigenvector is kept: the last rung
// of the robust-fallback ladder smears at >= 3000 K.
constexpr double kWideSmearing = 1000.0;
// Longest Broyden history whose normal equations fit the mixing kernel's LDS.
constexpr int kBroydenLdsHistory = 32;

// Whether an instance's next SCC iteration is the plain one of a captured
// bucket iteration: still iterating, no level-shift projector to apply, the
// eigenvector capacity enough for its smearing, a Broyden history that fits in
// LDS, and no eigensolver failure.  Every other iteration needs the host (level
// shift GEMMs, a larger eigenvector buffer, global Broyden scratch, an error).
// The device gate and the host loop evaluate this same function.
__host__ __device__ inline bool plain_iteration(const SccState& s, int kcap, int ld, int broyden_memory) {
    if (s.status != 0 || s.eig_fail || s.kneed_full > kcap) return false;
    if (s.have_prev_p && s.level_shift != 0.0) return false;
    if (s.temperature > kWideSmearing && kcap < ld) return false;
    const int m = s.broyden_iter < broyden_memory ? s.broyden_iter : broyden_memory;
    return m <= kBroydenLdsHistory;
}

struct DeviceTables {
    const ShellType* shells;     // (kMaxZ+1)*kMaxShellsPerAtom
    const ElementType* elements; // kMaxZ+1
    const double* d3_c6;         // D3 reference C6 (layout of data/make_data.py)
    const double* pair_scale;    // (kMaxZ+1)^2 H0 pair scaling k_AB
    Globals glob;
};

// Periodic cell of one instance (the Gamma-point path, pbc/mod.rs:83-91).
// Molecular instances have on = 0 and are never read beyond that flag.
struct CellRec {
    double a[3][3];    // cell vectors, one per row (bohr)
    double h[3];       // plane heights: the distance between lattice planes normal to a x b etc.
    double vol;        // |det|
    double alpha;      // Ewald splitting, sqrt(pi) / V^(1/3) (pbc/ewald.rs:48-60)
    double ew_self;    // -2 alpha / sqrt(pi), the 1/R self term of an atom
    double ew_bg;      // -pi / (alpha^2 V), the neutralising background per atom pair
    double r3_k0;      // -1/2 * 4 pi [ln k + (ln pi - psi(3/2)) / 2] / V, k = alpha / sqrt(pi), per eta^-2
    double r3_self;    // -1/2 * -(4 pi / 3) k^3, per eta^-2, same-atom pairs
    int64_t g0;        // first reciprocal vector of the instance in the G table
    int64_t gam0;      // offset of the instance's shell Coulomb matrix (nsh x nsh)
    int64_t ket0;      // offset of the instance's ket-share rows of the pair gradient (4 nat x nat)
    int32_t ng;        // reciprocal vectors with 0 < |G| <= 2 alpha tau
    int32_t on;        // periodic instance
};
// G table entry: G (bohr^-1), w1 = 4 pi / V exp(-G^2 / 4 alpha^2) / G^2 (Ewald
// 1/R), w3 = -1/2 * 2 pi / V * E1(G^2 / 4 alpha^2) (generalised Ewald R^-3).
constexpr int kGStride = 5;
// Charge structure factors per reciprocal vector, rewritten before every Ewald
// gradient: sum_j w_j cos(G.R_j) and sum_j w_j sin(G.R_j) over the shells j of
// the instance for the weights w = q_j, q_j / h_j and q_j / h_j^2 (h_j the shell
// hardness), cosine sums first.
constexpr int kGsfStride = 6;
// Cutoffs of the periodic path: AO image sums (pbc/mod.rs:170), the Ewald
// real-space remainder (:162), the Ewald convergence factor tau (pbc/ewald.rs:39).
constexpr double kPbcAoCutoff = 30.0;
constexpr double kEwaldShortCutoff = 10.0;
constexpr double kEwaldTau = 5.5;
// Largest periodic atom count for which the pair gradient evaluates each
// symmetry-unique pair once (its ket-share accumulator, 32 bytes per atom, sits
// in LDS); larger cells evaluate every pair from both sides.
constexpr int kKetShareMaxAtoms = 512;

// Flat (instance-major) index tables of a batch.
struct DeviceBatch {
    int32_t ninst;
    const int32_t* atom0;    // ninst+1
    const int32_t* sh0;      // ninst+1
    const int32_t* ao0;      // ninst+1
    const int32_t* ld;       // padded dimension per instance
    const int32_t* nsh;      // shells per instance (sh0 starts are padded)
    const int32_t* nao;      // AOs per instance (ao0 starts are padded)
    const int32_t* atom_inst;
    const int32_t* atom_z;
    const double* atom_x;    // 3 per atom
    const int32_t* atom_sh0; // global index of the atom's first shell
    const int32_t* sh_atom;  // global atom index
    const int32_t* sh_type;  // z*kMaxShellsPerAtom + s
    const int32_t* sh_ao;    // first AO, local to the instance
    const int32_t* ao_sh;    // global shell index
    const int32_t* sh_by_l;  // global shell indices, grouped by l within each instance
    const int32_t* lofs;     // 4 per instance: start of the l = 0, 1, 2 groups in sh_by_l, end
    const int64_t* vec0;     // offset of the per-instance length-ld vectors
    // per-instance dense matrices (column-major, ld x ld)
    double* const* S;
    double* const* H0;
    double* const* F;
    double* const* P;
    double* const* W;
    double* const* CB0;
    double* const* CB1;
    double* const* T;
    double* const* A;
    // periodic instances (nullptr when the batch has none)
    const CellRec* cell;     // ninst
    const double* gtab;      // kGStride per reciprocal vector
    const double* gam;       // shell Coulomb matrices of the periodic instances
    double* gsf;             // kGsfStride per reciprocal vector (gradient scratch)
};

__host__ __device__ inline bool periodic_instance(const DeviceBatch& b, int inst) {
    return b.cell != nullptr && b.cell[inst].on != 0;
}

}  // namespace gfn1
