// Grid-energy rescoring and rigid-body refinement of retained poses.
//
// The score of a pose with atoms at r_a (absolute coordinates, A) is
//   E = sum_a  sqrt(eps_a) * P_probe(a)(r_a) + q_a * P_elec(r_a),
// with P(r) the trilinear interpolation of a protein potential grid at
// u = (r - gridMin) / h.  Spreading an atom onto the 8 corners of its cell
// with the same weights and correlating (what the FFT search does,
// FillLigandGrid.cpp:41-66 and TranRotaConfSearchNoMini.cu:376-464) gives
// exactly this sum, so at a pose the search returns (pose rebuild of
// NoMini:468-487) the score equals the search's energy up to FP32 rounding.
// These kernels evaluate it in FP64, with its gradient, for poses handed to a
// minimiser (the grid forces of AddGridForcesToOpenMMSystem.cpp:54,100 use a
// spline interpolant instead, so they differ from this score by the
// interpolation).
#include <hip/hip_runtime.h>

#include <cmath>

#include "rescore.hpp"

namespace cdk {

namespace {

struct AtomTerm {
  double e;
  double g[3];
  int inside;
};

// Energy and gradient of one atom.  grid: [nch][N][N][N], z fastest.
__device__ __forceinline__ AtomTerm atom_term(const float* __restrict__ grid, int N, int pch, int ech, double wv,
                                              double wq, double x, double y, double z, double g0, double g1,
                                              double g2, double inv_h) {
  AtomTerm t{0.0, {0.0, 0.0, 0.0}, 0};
  const double u[3] = {(x - g0) * inv_h, (y - g1) * inv_h, (z - g2) * inv_h};
  int i[3];
  double f[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double fl = floor(u[k]);
    i[k] = (int)fl;
    f[k] = u[k] - fl;
    if (!(fl >= 0.0) || i[k] > N - 2) return t;  // outside (or NaN): no term
  }
  t.inside = 1;
  const size_t G = (size_t)N * N * N;
  const size_t base = ((size_t)i[0] * N + i[1]) * N + i[2];
  const float* pv = pch >= 0 ? grid + (size_t)pch * G + base : nullptr;
  const float* pe = grid + (size_t)ech * G + base;
  double e = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int dx = c >> 2, dy = (c >> 1) & 1, dz = c & 1;
    const size_t off = ((size_t)dx * N + dy) * N + dz;
    const double X = dx ? f[0] : 1.0 - f[0], Y = dy ? f[1] : 1.0 - f[1], Z = dz ? f[2] : 1.0 - f[2];
    const double sx = dx ? 1.0 : -1.0, sy = dy ? 1.0 : -1.0, sz = dz ? 1.0 : -1.0;
    double v = wq * (double)pe[off];
    if (pv) v += wv * (double)pv[off];
    e += X * Y * Z * v;
    d0 += sx * Y * Z * v;
    d1 += X * sy * Z * v;
    d2 += X * Y * sz * v;
  }
  t.e = e;
  t.g[0] = d0 * inv_h;
  t.g[1] = d1 * inv_h;
  t.g[2] = d2 * inv_h;
  return t;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// One wavefront per pose; lanes stride over atoms, then a fixed-order
// butterfly reduction (the same for every launch, so results are
// reproducible).
__global__ __launch_bounds__(256) void k_rescore(int npose, int natom, const double* __restrict__ coords,
                                                 const int* __restrict__ pch, const float* __restrict__ wv,
                                                 const float* __restrict__ wq, const float* __restrict__ grid, int N,
                                                 int ech, double g0, double g1, double g2, double inv_h,
                                                 double* __restrict__ energy, double* __restrict__ grad,
                                                 int* __restrict__ inside) {
  const int pose = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (pose >= npose) return;
  const double* r = coords + (size_t)pose * natom * 3;
  double e = 0.0;
  int n_in = 0;
  for (int a = lane; a < natom; a += 64) {
    const AtomTerm t = atom_term(grid, N, pch[a], ech, (double)wv[a], (double)wq[a], r[a * 3 + 0], r[a * 3 + 1],
                                 r[a * 3 + 2], g0, g1, g2, inv_h);
    e += t.e;
    n_in += t.inside;
    if (grad) {
      double* g = grad + ((size_t)pose * natom + a) * 3;
      g[0] = t.g[0];
      g[1] = t.g[1];
      g[2] = t.g[2];
    }
  }
  e = wave_sum(e);
  n_in = wave_sum_i(n_in);
  if (lane == 0) {
    energy[pose] = e;
    if (inside) inside[pose] = n_in;
  }
}

// Rigid-body descent on the grid score, one wavefront per pose.  Each step
// evaluates E, the force F = -sum g_a and the torque about the centroid
// tau = sum (r_a - c) x (-g_a); the trial move translates by st along F and
// rotates by sr about tau (Rodrigues).  A trial that lowers E + penalty is
// kept and both steps grow by 1.2 (capped), otherwise they halve.  Atoms off
// the grid add `penalty` each, so a pose cannot lower its score by leaving.
__device__ double pose_score(const double* __restrict__ r, int natom, const int* __restrict__ pch,
                             const float* __restrict__ wv, const float* __restrict__ wq,
                             const float* __restrict__ grid, int N, int ech, double g0, double g1, double g2,
                             double inv_h, double penalty, double* F, double* tau, const double* c) {
  const int lane = threadIdx.x & 63;
  double e = 0.0, f[3] = {0, 0, 0}, t[3] = {0, 0, 0};
  for (int a = lane; a < natom; a += 64) {
    const double x = r[a * 3 + 0], y = r[a * 3 + 1], z = r[a * 3 + 2];
    const AtomTerm at = atom_term(grid, N, pch[a], ech, (double)wv[a], (double)wq[a], x, y, z, g0, g1, g2, inv_h);
    e += at.inside ? at.e : penalty;
    const double fx = -at.g[0], fy = -at.g[1], fz = -at.g[2];
    f[0] += fx;
    f[1] += fy;
    f[2] += fz;
    const double dx = x - c[0], dy = y - c[1], dz = z - c[2];
    t[0] += dy * fz - dz * fy;
    t[1] += dz * fx - dx * fz;
    t[2] += dx * fy - dy * fx;
  }
  for (int k = 0; k < 3; ++k) {
    F[k] = wave_sum(f[k]);
    tau[k] = wave_sum(t[k]);
  }
  return wave_sum(e);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One 64-thread workgroup per pose; the pose and the trial pose live in LDS
// (2 * natom * 3 doubles), so every lane sees the others' writes in order.
__global__ __launch_bounds__(64) void k_refine(int npose, int natom, double* __restrict__ coords,
                                               const int* __restrict__ pch, const float* __restrict__ wv,
                                               const float* __restrict__ wq, const float* __restrict__ grid, int N,
                                               int ech, double g0, double g1, double g2, double inv_h, int steps,
                                               double penalty, double* __restrict__ energy) {
  extern __shared__ double sm[];
  const int pose = blockIdx.x;
  const int lane = threadIdx.x;
  if (pose >= npose) return;
  double* gr = coords + (size_t)pose * natom * 3;
  double* r = sm;
  double* q = sm + natom * 3;
  for (int i = lane; i < natom * 3; i += 64) r[i] = gr[i];
  wave_sync();
  double c[3];
  for (int k = 0; k < 3; ++k) {
    double s = 0.0;
    for (int a = lane; a < natom; a += 64) s += r[a * 3 + k];
    c[k] = wave_sum(s) / (double)natom;
  }
  double F[3], tau[3];
  double e = pose_score(r, natom, pch, wv, wq, grid, N, ech, g0, g1, g2, inv_h, penalty, F, tau, c);
  double st = 0.1, sr = 0.02;
  const double st_max = 0.5, sr_max = 0.1;
  for (int it = 0; it < steps; ++it) {
    const double fn = sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
    const double tn = sqrt(tau[0] * tau[0] + tau[1] * tau[1] + tau[2] * tau[2]);
    if (fn == 0.0 && tn == 0.0) break;
    double d[3], ax[3];
    for (int k = 0; k < 3; ++k) {
      d[k] = fn > 0.0 ? st * F[k] / fn : 0.0;
      ax[k] = tn > 0.0 ? tau[k] / tn : 0.0;
    }
    const double th = tn > 0.0 ? sr : 0.0;
    const double ct = cos(th), sn = sin(th);
    for (int a = lane; a < natom; a += 64) {
      const double v0 = r[a * 3 + 0] - c[0], v1 = r[a * 3 + 1] - c[1], v2 = r[a * 3 + 2] - c[2];
      const double kd = ax[0] * v0 + ax[1] * v1 + ax[2] * v2;
      const double cx = ax[1] * v2 - ax[2] * v1, cy = ax[2] * v0 - ax[0] * v2, cz = ax[0] * v1 - ax[1] * v0;
      q[a * 3 + 0] = c[0] + d[0] + v0 * ct + cx * sn + ax[0] * kd * (1.0 - ct);
      q[a * 3 + 1] = c[1] + d[1] + v1 * ct + cy * sn + ax[1] * kd * (1.0 - ct);
      q[a * 3 + 2] = c[2] + d[2] + v2 * ct + cz * sn + ax[2] * kd * (1.0 - ct);
    }
    wave_sync();
    const double cn[3] = {c[0] + d[0], c[1] + d[1], c[2] + d[2]};
    double F2[3], tau2[3];
    const double e2 = pose_score(q, natom, pch, wv, wq, grid, N, ech, g0, g1, g2, inv_h, penalty, F2, tau2, cn);
    if (e2 < e) {  // wave-uniform: every lane holds the reduced sums
      for (int i = lane; i < natom * 3; i += 64) r[i] = q[i];
      e = e2;
      for (int k = 0; k < 3; ++k) {
        c[k] = cn[k];
        F[k] = F2[k];
        tau[k] = tau2[k];
      }
      st = fmin(st * 1.2, st_max);
      sr = fmin(sr * 1.2, sr_max);
    } else {
      st *= 0.5;
      sr *= 0.5;
    }
    wave_sync();
  }
  for (int i = lane; i < natom * 3; i += 64) gr[i] = r[i];
  if (lane == 0) energy[pose] = e;
}

// CHARMM grid order (z fastest) to OpenMM's Continuous3DFunction order
// (x fastest), widened to double (AddGridForcesToOpenMMSystem.cpp:30-41).
__global__ __launch_bounds__(256) void k_grid_xfast(const float* __restrict__ in, int X, int Y, int Z,
                                                    double* __restrict__ out) {
  const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)X * Y * Z;
  if (o >= n) return;
  const int i = (int)(o % X);
  const int j = (int)((o / X) % Y);
  const int k = (int)(o / ((size_t)X * Y));
  out[o] = (double)in[((size_t)i * Y + j) * Z + k];
}

}  // namespace

hipError_t launch_grid_xfast(hipStream_t st, const float* in, int X, int Y, int Z, double* out) {
  const size_t n = (size_t)X * Y * Z;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_grid_xfast, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, X, Y, Z, out);
  return hipGetLastError();
}

hipError_t launch_rescore(hipStream_t st, int npose, int natom, const double* coords, const int* pch,
                          const float* wv, const float* wq, const float* grid, int N, int ech, const double* gmin,
                          double h, double* energy, double* grad, int* inside) {
  if (npose <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((npose + 3) / 4);
  hipLaunchKernelGGL(k_rescore, dim3(blocks), dim3(256), 0, st, npose, natom, coords, pch, wv, wq, grid, N, ech,
                     gmin[0], gmin[1], gmin[2], 1.0 / h, energy, grad, inside);
  return hipGetLastError();
}

hipError_t launch_refine(hipStream_t st, int npose, int natom, double* coords, const int* pch, const float* wv,
                         const float* wq, const float* grid, int N, int ech, const double* gmin, double h, int steps,
                         double penalty, double* energy) {
  if (npose <= 0) return hipSuccess;
  const size_t lds = (size_t)natom * 6 * sizeof(double);
  hipLaunchKernelGGL(k_refine, dim3((unsigned)npose), dim3(64), lds, st, npose, natom, coords, pch, wv, wq, grid, N,
                     ech, gmin[0], gmin[1], gmin[2], 1.0 / h, steps, penalty, energy);
  return hipGetLastError();
}

}  // namespace cdk
