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
};

}  // namespace gfn1
