// Protein potential grids built on the device from the protein's atoms.
//
// Channel p < nprobe (soft-core Lennard-Jones for a probe of radius r_p):
//   sum_i sqrt(eps_i) * ((R/d)^12 - 2 (R/d)^6),  R = r_i + r_p,
// over atoms with d <= lj_cutoff, then capped at +lj_cap.
// Channel nprobe (Coulomb with a distance-dependent dielectric 4d):
//   sum_i 332.0636 q_i / (4 d^2) over atoms with d <= elec_cutoff, clamped
// to [-elec_cap, elec_cap].  d^2 is floored at 1e-4 A^2.
// Grid point (i, j, k) sits at gmin + (i, j, k) * h.  Each point sums its
// atoms in ascending index order in FP64 with no contraction, so the result
// is fixed by the inputs (independent of the launch shape).
#include <hip/hip_runtime.h>

#include "gridbuild.hpp"

namespace cdk {

namespace {

constexpr int kTile = 256;     // atoms staged in LDS per pass
constexpr int kMaxProbe = 15;  // probe channels

__global__ __launch_bounds__(256) void k_build_grids(int natom, const double* __restrict__ atoms, int nprobe,
                                                     const double* __restrict__ probe, int n, double g0, double g1,
                                                     double g2, double h, double rc_lj2, double rc_el2,
                                                     double lj_cap, double el_cap, float* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double sx[kTile], sy[kTile], sz[kTile], sq[kTile], sr[kTile], se[kTile];
  const size_t G = (size_t)n * n * n;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = g < G;
  const int ix = (int)(g / ((size_t)n * n)), iy = (int)((g / n) % n), iz = (int)(g % n);
  const double gx = g0 + ix * h, gy = g1 + iy * h, gz = g2 + iz * h;
  double lj[kMaxProbe];
  double pr[kMaxProbe];
#pragma unroll
  for (int p = 0; p < kMaxProbe; ++p) {
    lj[p] = 0.0;
    pr[p] = p < nprobe ? probe[p] : 0.0;
  }
  double el = 0.0;
  for (int t0 = 0; t0 < natom; t0 += kTile) {
    __syncthreads();
    for (int t = threadIdx.x; t < kTile && t0 + t < natom; t += blockDim.x) {
      const double* a = atoms + (size_t)(t0 + t) * 6;
      sx[t] = a[0];
      sy[t] = a[1];
      sz[t] = a[2];
      sq[t] = a[3];
      sr[t] = a[4];
      se[t] = sqrt(a[5]);
    }
    __syncthreads();
    if (!live) continue;
    const int m = min(kTile, natom - t0);
    for (int t = 0; t < m; ++t) {
      const double dx = sx[t] - gx, dy = sy[t] - gy, dz = sz[t] - gz;
      double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 > rc_el2) continue;
      d2 = fmax(d2, 1e-4);
      el += 332.0636 * sq[t] / (4.0 * d2);
      if (d2 > rc_lj2) continue;
      const double inv2 = 1.0 / d2;
#pragma unroll
      for (int p = 0; p < kMaxProbe; ++p) {
        if (p < nprobe) {
          const double R = sr[t] + pr[p];
          const double s2 = R * R * inv2;
          const double s6 = s2 * s2 * s2;
          lj[p] += se[t] * (s6 * s6 - 2.0 * s6);
        }
      }
    }
  }
  if (!live) return;
#pragma unroll
  for (int p = 0; p < kMaxProbe; ++p)
    if (p < nprobe) out[(size_t)p * G + g] = (float)fmin(lj[p], lj_cap);
  out[(size_t)nprobe * G + g] = (float)fmax(-el_cap, fmin(el, el_cap));
}

}  // namespace

int grid_build_max_probes() { return kMaxProbe; }

hipError_t launch_build_grids(hipStream_t st, int natom, const double* atoms, int nprobe, const double* probe, int n,
                              const double* gmin, double h, double lj_cutoff, double elec_cutoff, double lj_cap,
                              double elec_cap, float* out) {
  if (nprobe < 0 || nprobe > kMaxProbe) return hipErrorInvalidValue;
  const size_t G = (size_t)n * n * n;
  const unsigned blocks = (unsigned)((G + 255) / 256);
  hipLaunchKernelGGL(k_build_grids, dim3(blocks), dim3(256), 0, st, natom, atoms, nprobe, probe, n, gmin[0], gmin[1],
                     gmin[2], h, lj_cutoff * lj_cutoff, elec_cutoff * elec_cutoff, lj_cap, elec_cap, out);
  return hipGetLastError();
}

}  // namespace cdk
