// Batched L-BFGS geometry optimisation over the batch engine.
#pragma once

#include <string>
#include <vector>

#include "engine.hpp"

namespace gfn1 {

// GeometryOptimizationOptions (src/optimizer.rs:18-45) with the same defaults
// (src/optimizer.rs:10-15).
struct OptimizerOptions {
    int max_iterations = 250;
    double gradient_tolerance = 1.0e-4;
    double step_tolerance = 1.0e-7;
    int history = 12;
    double initial_step = 1.0;
    double max_atom_step = 0.30;
    // 0: every evaluation starts its SCC as the batch was configured (the
    // reference's behaviour); 1: from the shell charges of the last accepted
    // geometry of the same system
    int warm_start = 0;
    int keep_trajectory = 1;
};

// GeometryOptimizationStep (src/optimizer.rs:47-55).
struct OptimizerStep {
    int iteration = 0;
    double energy = 0.0, max_gradient = 0.0, step_norm = 0.0;
    std::vector<double> positions;  // bohr, 3 * nat
};

// GeometryOptimizationResult (src/optimizer.rs:57-67) of one system, plus its
// status: kOk, or the error that optimize_geometry would have returned.
struct OptimizerResult {
    int status = kOk;
    std::string message;
    double energy = 0.0, max_gradient = 0.0;
    int iterations = 0;
    bool converged = false;
    std::vector<double> positions, gradient;  // bohr, Hartree/bohr; 3 * nat
    std::vector<OptimizerStep> trajectory;
};

struct OptimizerStats {
    int rounds = 0;             // batch runs
    long long evaluations = 0;  // energy + gradient evaluations over all systems
    double total_ms = 0.0;
};

// Runs optimize_geometry (src/optimizer.rs:78-185) for every instance of the
// batch at once.  Each system keeps its own L-BFGS state; every round
// evaluates the pending geometries of all systems (the start point or the
// next line-search trial) in one batch run over just those systems.
Status optimize_batch(Batch& batch, const OptimizerOptions& opt, std::vector<OptimizerResult>& out,
                      OptimizerStats* stats);

// Pieces of the iteration (exposed for tests of the host logic).
std::vector<double> lbfgs_direction(const std::vector<double>& g, const std::vector<std::vector<double>>& s,
                                    const std::vector<std::vector<double>>& y, const std::vector<double>& rho);
void limit_atom_step(std::vector<double>& d, double max_atom_step);
double max_gradient(const std::vector<double>& g);

}  // namespace gfn1
