// Random-torsion conformers on the device (conformers.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace cdk {

// U[0, 1) from a counter-based stream; the same on host and device.
__host__ __device__ double torsion_uniform(unsigned long long seed, int trial, int rotor);

// coords [natom][3]; rotors [nrotor][4]; moving [nrotor][natom] (0/1);
// exclude [natom][natom] (0/1, pairs the steric filter skips); out
// [ntrial][natom][3]; ok [ntrial] (1 = no clash).
hipError_t launch_conformers(hipStream_t st, int natom, const double* coords, int nrotor, const int* rotors,
                             const unsigned char* moving, const unsigned char* exclude, int ntrial,
                             unsigned long long seed, double clash_dist, double* out, int* ok);

}  // namespace cdk
