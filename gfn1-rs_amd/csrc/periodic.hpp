// Periodic Gamma-point path: Ewald shell Coulomb matrix, its gradient, and the
// host-side cell records (periodic.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "device.hpp"

namespace gfn1 {

// Cell record of one periodic instance from its three cell vectors (bohr, row
// after row), with its reciprocal vectors appended to gtab (kGStride doubles
// each).  False for a non-finite or singular cell (|det| < 1e-14).
bool make_cell_record(const double* cell, CellRec& c, std::vector<double>& gtab);
// E1(x) for x > 0.
double exponential_integral_e1(double x);

// Gamma (nsh x nsh) of the listed periodic instances into b.gam.
hipError_t launch_ewald_gamma(hipStream_t s, const DeviceBatch& b, const DeviceTables& t, const int* pinst,
                              int npinst, int max_nat);
// g_ew[3 a + d] = d(1/2 q^T Gamma q)/dR_a for the listed atoms of the listed
// periodic instances (max_ng: the largest G count among them); the charge
// structure factors go to b.gsf first.
hipError_t launch_ewald_grad(hipStream_t s, const DeviceBatch& b, const DeviceTables& t, const int* pinst,
                             int npinst, int max_ng, const int* patoms, int npatoms, const double* q, double* g_ew);

}  // namespace gfn1
