// Learned SCC start charges (guess.hip).
//
// A per-atom MLP maps a radial-environment descriptor of each atom to its two
// shell charges; the per-instance predictions are shifted so that they sum to
// the instance charge and are used as the SCC start vector (the reference's
// scc_initial_shell_charges hook, electronic.rs:386-392, 1231-1239).  The
// converged SCC does not depend on the start vector, so the guess only changes
// the number of iterations.
//
// Descriptor (68 values): one-hot of the element over {H, C, N, O}, then for
// each neighbour element 16 Gaussians exp(-eta (r - mu_k)^2) * fc(r), centres
// mu_k = 0.5 .. 5.0 A, eta = 1 / (2 * 0.3^2) A^-2, cosine cutoff fc at 5 A.
// Network: 68 -> 64 SiLU -> 64 SiLU -> 2, FP32, inputs standardised with the
// training-set mean and spread.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "model.hpp"

namespace gfn1 {

constexpr int kGuessFeat = 68;
constexpr int kGuessHidden = 64;
constexpr int kGuessOut = 2;
// parameter vector layout (row-major weights, output index first)
constexpr int kGoW1 = 0;
constexpr int kGoB1 = kGoW1 + kGuessHidden * kGuessFeat;
constexpr int kGoW2 = kGoB1 + kGuessHidden;
constexpr int kGoB2 = kGoW2 + kGuessHidden * kGuessHidden;
constexpr int kGoW3 = kGoB2 + kGuessHidden;
constexpr int kGoB3 = kGoW3 + kGuessOut * kGuessHidden;
constexpr int kGuessParams = kGoB3 + kGuessOut;  // 8706
// Limits of the no-allocation prediction (its device scratch is sized for them).
constexpr int kNoAllocMaxAtoms = 16384;
constexpr int kNoAllocMaxSystems = 16384;

// Element slot of the descriptor (0..3 for H, C, N, O), -1 otherwise.
inline int guess_element_slot(int z) { return z == 1 ? 0 : z == 6 ? 1 : z == 7 ? 2 : z == 8 ? 3 : -1; }

struct GuessTrainOptions {
    int epochs = 40;
    int batch_atoms = 1024;
    float learning_rate = 1e-3f;
    uint64_t seed = 1;
};

// Concatenated atoms of many systems (positions in bohr).
struct GuessAtoms {
    std::vector<int> z, inst, inst_atom0;  // inst_atom0 has n_inst + 1 entries
    std::vector<double> x;
    std::vector<int> usable;               // 1 if every atom of the instance has a descriptor slot
};

// Device kernels shared with the batch engine.
hipError_t guess_launch_features(hipStream_t s, int natoms, const int* z, const double* x, const int* inst,
                                 const int* inst_atom0, float* feat);
hipError_t guess_launch_forward(hipStream_t s, int natoms, const float* theta, const float* norm, const float* feat,
                                float* out);
hipError_t guess_launch_apply(hipStream_t s, int ninst, const int* usable, const int* atom0, const int* atom_sh0,
                              const double* charge, const float* pred, double* q);

class GuessModel {
   public:
    GuessModel() = default;
    ~GuessModel();
    GuessModel(const GuessModel&) = delete;
    GuessModel& operator=(const GuessModel&) = delete;

    Status init_random(uint64_t seed);
    Status load(const std::string& path);
    Status save(const std::string& path) const;
    // host copies: parameters then the standardisation (mean[68], inv_spread[68])
    const std::vector<float>& params() const { return theta_; }
    const std::vector<float>& norm() const { return norm_; }
    Status set_params(const float* theta);
    // device copies (uploaded on demand to the current device)
    Status device_ptrs(const float** theta, const float** norm) const;

    // Train on converged shell charges; labels are two per atom (shell order of
    // the atom).  loss_per_epoch receives opt.epochs mean-squared errors.
    Status train(const GuessAtoms& atoms, const std::vector<float>& labels, const GuessTrainOptions& opt,
                 std::vector<float>& loss_per_epoch);
    // Half mean squared error and its gradient over all usable atoms (full batch).
    Status loss_gradient(const GuessAtoms& atoms, const std::vector<float>& labels, double* loss,
                         std::vector<float>& grad) const;
    // Per-atom network outputs (two per atom, FP32) for usable atoms; zeros elsewhere.
    Status predict_atoms(const GuessAtoms& atoms, std::vector<float>& out) const;
    // The same outputs without allocating device memory: the model must already
    // be resident (it is, once created or loaded on a machine with a device), the
    // inputs at most kNoAllocMaxAtoms atoms in kNoAllocMaxSystems systems.  Calls
    // in one process share one device scratch array and are serialised.
    Status predict_atoms_noalloc(const GuessAtoms& atoms, std::vector<float>& out) const;

   private:
    Status upload() const;
    Status upload_if_device() const;
    std::vector<float> theta_, norm_;
    mutable float* d_theta_ = nullptr;
    mutable float* d_norm_ = nullptr;
    mutable bool dirty_ = true;
};

// Bytes of the static device scratch of predict_atoms_noalloc.
size_t guess_noalloc_scratch_bytes();

// Descriptors on the device for a set of atoms (natoms x kGuessFeat, FP32).
Status guess_features(const GuessAtoms& atoms, std::vector<float>& out);

}  // namespace gfn1
