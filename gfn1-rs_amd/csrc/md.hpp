// Batched velocity-Verlet molecular dynamics over the batch engine.
#pragma once

#include <vector>

#include "engine.hpp"

namespace gfn1 {

struct MdOptions {
    double timestep_fs = 0.5;
    int steps = 10;
    int warm_start = 1;  // each SCC starts from the previous step's converged charges
};

// NVE dynamics of every instance from the batch's current positions and the
// given velocities (3 per atom, bohr per atomic time unit; updated in place).
// pot / kin: (steps + 1) x size() energies (Eh), row = step; status: per
// instance, kOk or the error that stopped it (a stopped system stays put);
// iters: (steps + 1) x size() SCC iteration counts (-1 once a system stopped).
Status md_batch(Batch& batch, const MdOptions& opt, std::vector<double>& v, std::vector<double>& pot,
                std::vector<double>& kin, std::vector<int>& status, std::vector<int>& iters);

}  // namespace gfn1
