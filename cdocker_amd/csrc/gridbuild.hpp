// Protein potential grids built on the device from atoms (gridbuild.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace cdk {

int grid_build_max_probes();
// atoms [natom][6] = x, y, z, charge, radius, epsilon (A, e, A, kcal/mol);
// probe [nprobe] radii; out [nprobe + 1][n][n][n] (last = Coulomb).
hipError_t launch_build_grids(hipStream_t st, int natom, const double* atoms, int nprobe, const double* probe, int n,
                              const double* gmin, double h, double lj_cutoff, double elec_cutoff, double lj_cap,
                              double elec_cap, float* out);

}  // namespace cdk
