// Random-torsion conformers on the device.
//
// GeneRandomConformations.cpp:36-47 sets every rotatable bond of each of
// 2000 copies of the ligand to an angle drawn uniformly in [-pi, pi), then
// minimises each copy with OpenMM and keeps those under an energy cap
// (:49-109).  Here each trial sets the rotors in order to angles from a
// counter-based stream (seed, trial, rotor), keeps the trial when no
// non-excluded atom pair comes closer than clash_dist (a steric filter in
// place of the force-field energy, which needs OpenMM), and reports one flag
// per trial; the host keeps the first survivors in trial order.
//
// Setting a rotor (a, b, c, d) to angle theta: phi = the current (IUPAC)
// dihedral of a-b-c-d; the atoms on the d side of bond b-c turn about the
// axis b->c through c by theta - phi (Rodrigues, right-handed).  All in FP64
// with no contraction.
#include <hip/hip_runtime.h>

#include "conformers.hpp"

namespace cdk {

// splitmix64 finaliser on a (seed, trial, rotor) counter -> U[0, 1), 53 bits.
__host__ __device__ double torsion_uniform(unsigned long long seed, int trial, int rotor) {
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * ((unsigned long long)trial * 0x100000001ull +
                                                          (unsigned long long)rotor + 1ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

namespace {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One 64-thread workgroup per trial; the trial's coordinates live in LDS.
__global__ __launch_bounds__(64) void k_conformers(int natom, const double* __restrict__ coords, int nrotor,
                                                   const int* __restrict__ rotors,
                                                   const unsigned char* __restrict__ moving,
                                                   const unsigned char* __restrict__ exclude, int ntrial,
                                                   unsigned long long seed, double clash2,
                                                   double* __restrict__ out, int* __restrict__ ok) {
#pragma clang fp contract(off)
  extern __shared__ double sm[];
  const int t = blockIdx.x;
  const int lane = threadIdx.x;
  if (t >= ntrial) return;
  double* x = sm;
  for (int i = lane; i < natom * 3; i += 64) x[i] = coords[i];
  wave_sync();
  const double kPi = 3.14159265358979323846;
  for (int r = 0; r < nrotor; ++r) {
    const int a = rotors[r * 4 + 0], b = rotors[r * 4 + 1], c = rotors[r * 4 + 2], d = rotors[r * 4 + 3];
    // current dihedral (every lane computes the same value)
    const double b1x = x[b * 3 + 0] - x[a * 3 + 0], b1y = x[b * 3 + 1] - x[a * 3 + 1], b1z = x[b * 3 + 2] - x[a * 3 + 2];
    const double b2x = x[c * 3 + 0] - x[b * 3 + 0], b2y = x[c * 3 + 1] - x[b * 3 + 1], b2z = x[c * 3 + 2] - x[b * 3 + 2];
    const double b3x = x[d * 3 + 0] - x[c * 3 + 0], b3y = x[d * 3 + 1] - x[c * 3 + 1], b3z = x[d * 3 + 2] - x[c * 3 + 2];
    const double n1x = b1y * b2z - b1z * b2y, n1y = b1z * b2x - b1x * b2z, n1z = b1x * b2y - b1y * b2x;
    const double n2x = b2y * b3z - b2z * b3y, n2y = b2z * b3x - b2x * b3z, n2z = b2x * b3y - b2y * b3x;
    const double l2 = sqrt(b2x * b2x + b2y * b2y + b2z * b2z);
    // phi = atan2(b2^ . (n1 x n2), n1 . n2), the IUPAC dihedral
    const double qx = n1y * n2z - n1z * n2y, qy = n1z * n2x - n1x * n2z, qz = n1x * n2y - n1y * n2x;
    const double yv = (qx * b2x + qy * b2y + qz * b2z) / l2;
    const double xv = n1x * n2x + n1y * n2y + n1z * n2z;
    const double phi = atan2(yv, xv);
    const double theta = -kPi + 2.0 * kPi * torsion_uniform(seed, t, r);
    const double ang = theta - phi;
    const double ux = b2x / l2, uy = b2y / l2, uz = b2z / l2;
    const double ca = cos(ang), sa = sin(ang);
    const double cx = x[c * 3 + 0], cy = x[c * 3 + 1], cz = x[c * 3 + 2];
    const unsigned char* mv = moving + (size_t)r * natom;
    wave_sync();  // every lane has read the rotor's atoms
    for (int i = lane; i < natom; i += 64) {
      if (!mv[i]) continue;
      const double vx = x[i * 3 + 0] - cx, vy = x[i * 3 + 1] - cy, vz = x[i * 3 + 2] - cz;
      const double kd = ux * vx + uy * vy + uz * vz;
      const double kx = uy * vz - uz * vy, ky = uz * vx - ux * vz, kz = ux * vy - uy * vx;
      x[i * 3 + 0] = cx + vx * ca + kx * sa + ux * kd * (1.0 - ca);
      x[i * 3 + 1] = cy + vy * ca + ky * sa + uy * kd * (1.0 - ca);
      x[i * 3 + 2] = cz + vz * ca + kz * sa + uz * kd * (1.0 - ca);
    }
    wave_sync();
  }
  // steric filter over the non-excluded pairs
  int bad = 0;
  for (int i = lane; i < natom; i += 64)
    for (int j = i + 1; j < natom; ++j) {
      if (exclude[(size_t)i * natom + j]) continue;
      const double dx = x[i * 3 + 0] - x[j * 3 + 0], dy = x[i * 3 + 1] - x[j * 3 + 1], dz = x[i * 3 + 2] - x[j * 3 + 2];
      if (dx * dx + dy * dy + dz * dz < clash2) bad = 1;
    }
  const unsigned long long anybad = __ballot(bad);
  for (int i = lane; i < natom * 3; i += 64) out[(size_t)t * natom * 3 + i] = x[i];
  if (lane == 0) ok[t] = anybad ? 0 : 1;
}

}  // namespace

hipError_t launch_conformers(hipStream_t st, int natom, const double* coords, int nrotor, const int* rotors,
                             const unsigned char* moving, const unsigned char* exclude, int ntrial,
                             unsigned long long seed, double clash_dist, double* out, int* ok) {
  if (ntrial <= 0) return hipSuccess;
  const size_t lds = (size_t)natom * 3 * sizeof(double);
  hipLaunchKernelGGL(k_conformers, dim3((unsigned)ntrial), dim3(64), lds, st, natom, coords, nrotor, rotors, moving,
                     exclude, ntrial, seed, clash_dist * clash_dist, out, ok);
  return hipGetLastError();
}

}  // namespace cdk
