// Cartesian Hessians from central differences of the analytic gradient, with
// every displaced geometry of every system evaluated in batched device passes.
#pragma once

#include <string>
#include <vector>

#include "engine.hpp"

namespace gfn1 {

struct HessianOptions {
    double step = 1.0e-4;                 // bohr, central differences
    double scc_energy_tolerance = 1.0e-11;  // SCC tolerances of every evaluation
    double scc_charge_tolerance = 1.0e-11;
    // gradient terms differentiated (AnalyticHessianOptions, src/hessian/mod.rs:28-52;
    // include_electronic stands for its four electronic flags together)
    int include_repulsion = 1, include_electronic = 1, include_dispersion = 1, include_halogen = 1;
    // displaced SCCs start from the converged shell charges of the undisplaced
    // geometry: 5-6 SCC iterations instead of 9-10.  Such an SCC stops closer
    // to its start, so the charge tolerance is 1e-11: with 1e-10 the Hessians
    // of methane and ammonia were off by 9e-8, with 1e-11 by at most 7e-9
    int warm_start = 1;
    int max_batch = 4096;  // displaced geometries per device pass
};

struct HessianResult {
    int status = kOk;
    std::string message;
    double energy = 0.0;           // total free energy at the undisplaced geometry
    std::vector<double> gradient;  // 3 nat, undisplaced geometry
    std::vector<double> hessian;   // (3 nat)^2, row-major, symmetrised
    double max_asymmetry = 0.0;    // max |H_ij - H_ji| before symmetrisation
    double dipole[3] = {0.0, 0.0, 0.0};   // Mulliken dipole (e bohr), undisplaced geometry
    std::vector<double> dipole_derivatives;  // d mu_alpha / d R_c at [3 c + alpha], central differences
    // the repulsion, D3 and halogen parts of the Hessian ((3 nat)^2 each, same
    // layout and differences, symmetrised); the electronic part is the rest
    std::vector<double> h_repulsion, h_dispersion, h_halogen;
};

struct HessianStats {
    long long evaluations = 0;
    int passes = 0;
    double total_ms = 0.0;
};

Status hessian_batch(const DeviceModel& dm, const std::vector<SystemInput>& systems, EngineOptions e,
                     const HessianOptions& opt, std::vector<HessianResult>& out, HessianStats* stats);

}  // namespace gfn1
