// gfx950 kernels of the FFT translational-rotational pose search.
//
// Two pipelines share these kernels:
//   * the reference-dataflow ("plain") pipeline: full-grid spread -> batched
//     3-D R2C -> ConjMult -> SumGrids -> 3-D C2R -> windowed argmin, one
//     materialised array per stage, exactly the stage boundaries of
//     TranRotaConfSearchNoMini.cu:389-464;
//   * the fused, pruned pipeline: prep -> (spread + z/y forward FFT of the
//     occupied x-planes only) -> (x FFT + conj-multiply + channel sum +
//     inverse x FFT, protein spectra staged in LDS and reused across poses)
//     -> (inverse y FFT + z C2R of the window rows only + argmin).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <type_traits>

#include "cdk_types.hpp"
#include "fft_lds.hpp"
#include "fft_reg.hpp"
#include "kernels.hpp"
#include "xcd_map.hpp"

namespace cdk {

// ---------------------------------------------------------------------------
// Plans per grid edge.  The tuned radix list is selected at run time among
// the variants compiled here (see tune_*).
// ---------------------------------------------------------------------------
using P32a = Plan<32, 8, 4>;
using P48a = Plan<48, 16, 3>;
using P48b = Plan<48, 4, 4, 3>;
using P64a = Plan<64, 8, 8>;
using P64b = Plan<64, 4, 4, 4>;
using P96a = Plan<96, 8, 4, 3>;
using P96b = Plan<96, 16, 2, 3>;
using P128a = Plan<128, 8, 16>;
using P128b = Plan<128, 4, 4, 8>;

// ---------------------------------------------------------------------------
// Ordered 64-bit argmin key: high word = float bits mapped to unsigned order,
// low word = linear translation index.  min(key) = lowest energy, and among
// equal energies the lowest (i, j, k) in x-major order, which is the first
// strict minimum of the reference's scan (TranRotaConfSearchNoMini.cu:441-464).
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long energy_key(float e, unsigned lin) {
  unsigned u = __float_as_uint(e);
  u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return ((unsigned long long)u << 32) | (unsigned long long)lin;
}

// The key's high word alone: u ^ (sign ? ~0 : 0x80000000).
__device__ __forceinline__ unsigned energy_bits(float e) {
  const unsigned u = __float_as_uint(e);
  return u ^ ((unsigned)((int)u >> 31) | 0x80000000u);
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned lo = __shfl_xor((unsigned)(v & 0xffffffffu), off, 64);
    unsigned hi = __shfl_xor((unsigned)(v >> 32), off, 64);
    unsigned long long o = ((unsigned long long)hi << 32) | lo;
    v = o < v ? o : v;
  }
  return v;
}

// Block-wide min of a key, then one atomicMin into `dst` (order independent,
// so the result does not depend on workgroup scheduling).  At most NW
// wavefronts per workgroup.
template <int NW = kBlock / 64>
__device__ __forceinline__ void block_min_commit(unsigned long long v, unsigned long long* dst) {
  __shared__ unsigned long long red[NW];
  v = wave_min_u64(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = red[i] < m ? red[i] : m;
    if (m != ~0ull) atomicMin(dst, m);
  }
}

// ---------------------------------------------------------------------------
// Raw buffer access.  A descriptor covers [base, base + bytes); loads past the
// end return zero and stores past the end are dropped, which is how the
// kernels zero-pad the ligand box and clip the translation window without
// per-element compares.  The descriptor is wave-uniform (scalar registers).
// ---------------------------------------------------------------------------
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float2 buffer_load_f2(__amdgpu_buffer_rsrc_t r, int voff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
}
__device__ __forceinline__ float4 buffer_load_f4(__amdgpu_buffer_rsrc_t r, int voff) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0));
}
__device__ __forceinline__ void buffer_store_f2(float2 v, __amdgpu_buffer_rsrc_t r, int voff) {
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, 0, 0);
}
// Offset that is past every descriptor (lanes with no line to serve).
constexpr int kVoffNone = 0x40000000;

// ---------------------------------------------------------------------------
// Rotate.cpp:4-27 with QuaternionMultiply.cpp:3-11 term order, FP32, no FMA
// contraction, so the rotated coordinates are bit-identical to the host code.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void rotate_exact(const float* q, float px, float py, float pz, float* o) {
#pragma clang fp contract(off)
  const float zero = 0.0f;
  const float t0 = q[0] * zero - q[1] * px - q[2] * py - q[3] * pz;
  const float t1 = q[0] * px + q[1] * zero + q[2] * pz - q[3] * py;
  const float t2 = q[0] * py - q[1] * pz + q[2] * zero + q[3] * px;
  const float t3 = q[0] * pz + q[1] * py - q[2] * px + q[3] * zero;
  const float c0 = q[0], c1 = -q[1], c2 = -q[2], c3 = -q[3];
  o[0] = t0 * c1 + t1 * c0 + t2 * c3 - t3 * c2;
  o[1] = t0 * c2 - t1 * c3 + t2 * c0 + t3 * c1;
  o[2] = t0 * c3 + t1 * c2 - t2 * c1 + t3 * c0;
}

// FillLigandGrid.cpp:24-35: grid index by truncation and the double ratio,
// where `xIndex * spacing` is a float product.
__device__ __forceinline__ void grid_place(float x, float mn, float h, int& idx, double& f) {
#pragma clang fp contract(off)
  const double d = (double)x - (double)mn;
  idx = (int)(d / (double)h);
  const float ih = (float)idx * h;
  f = (d - (double)ih) / (double)h;
}

// (value, index) lexicographic min/max, reproducing the first-occurrence
// semantics of GetMinCoors.cpp / GetMaxCoors.cpp (strict comparisons).
__device__ __forceinline__ void lex_min(float& v, int& i, float ov, int oi) {
  if (ov < v || (ov == v && oi < i)) { v = ov; i = oi; }
}
__device__ __forceinline__ void lex_max(float& v, int& i, float ov, int oi) {
  if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
}

// One wavefront per pose: rotate, AABB, window, per-atom grid placement.
__global__ __launch_bounds__(kPrepBlock) void k_prep(const PoseSlot* __restrict__ slots, int nslots,
                                                     const LigandDesc* __restrict__ ligs,
                                                     const float* __restrict__ coords, int N, float h, int astride,
                                                     int boxmax, float* __restrict__ rot, AtomGrid* __restrict__ ag,
                                                     RotMeta* __restrict__ meta,
                                                     unsigned long long* __restrict__ keys) {
  const int s = blockIdx.x;
  if (s >= nslots) return;
  const PoseSlot ps = slots[s];
  const int na = ligs[ps.lig].natom;
  const float* c = coords + ps.coord_off;
  const int lane = threadIdx.x;
  float mn[3], mx[3];
  int mni[3], mxi[3];
  for (int k = 0; k < 3; ++k) {
    mn[k] = __int_as_float(0x7f800000);
    mx[k] = -__int_as_float(0x7f800000);
    mni[k] = INT_MAX;
    mxi[k] = INT_MAX;
  }
  for (int a = lane; a < na; a += kPrepBlock) {
    float o[3];
    rotate_exact(ps.q, c[a * 3 + 0], c[a * 3 + 1], c[a * 3 + 2], o);
    float* r = rot + ((size_t)s * astride + a) * 3;
    r[0] = o[0];
    r[1] = o[1];
    r[2] = o[2];
    for (int k = 0; k < 3; ++k) {
      lex_min(mn[k], mni[k], o[k], a);
      lex_max(mx[k], mxi[k], o[k], a);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    for (int k = 0; k < 3; ++k) {
      float ov = __shfl_xor(mn[k], off, 64);
      int oi = __shfl_xor(mni[k], off, 64);
      lex_min(mn[k], mni[k], ov, oi);
      ov = __shfl_xor(mx[k], off, 64);
      oi = __shfl_xor(mxi[k], off, 64);
      lex_max(mx[k], mxi[k], ov, oi);
    }
  }
  // Every lane now holds the pose's min/max.  Place atoms (needs the min).
  int bmax[3] = {0, 0, 0};
  for (int a = lane; a < na; a += kPrepBlock) {
    const float* r = rot + ((size_t)s * astride + a) * 3;
    AtomGrid g;
    grid_place(r[0], mn[0], h, g.ix, g.fx);
    grid_place(r[1], mn[1], h, g.iy, g.fy);
    grid_place(r[2], mn[2], h, g.iz, g.fz);
    g.pad = 0;
    ag[(size_t)s * astride + a] = g;
    bmax[0] = max(bmax[0], g.ix);
    bmax[1] = max(bmax[1], g.iy);
    bmax[2] = max(bmax[2], g.iz);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 3; ++k) bmax[k] = max(bmax[k], __shfl_xor(bmax[k], off, 64));
  if (lane == 0) {
    RotMeta m;
    int ok = na > 0;
    for (int k = 0; k < 3; ++k) {
      m.minc[k] = mn[k];
      const float len = mx[k] - mn[k];
      m.len[k] = len;
      m.box[k] = bmax[k] + 2;
      m.win[k] = N - (int)__fdiv_rn(len, h) - 2;
      ok = ok && m.win[k] > 0 && m.box[k] <= N && m.box[k] <= boxmax;
    }
    m.valid = ok;
    m.pad = 0;
    meta[s] = m;
    keys[s] = ~0ull;
  }
}

// Per-atom weight of ligand channel c: charge for the last (elec) channel,
// sqrt(epsilon) otherwise (FillLigandGrid.cpp:45, :97).
__device__ __forceinline__ float channel_weight(const LigandDesc& L, int c, int a, const float* wv,
                                                const float* wq) {
  return (c == L.nch - 1) ? wq[L.w_off + a] : wv[L.w_off + a];
}

// One trilinear contribution, FillLigandGrid.cpp:45-66:
//   w * X * Y * Z evaluated left to right in double, added to the float cell.
__device__ __forceinline__ float accum_corner(float acc, double wX, double fy, double fz, int dy, int dz) {
#pragma clang fp contract(off)
  const double Y = dy ? fy : (1.0 - fy);
  const double Z = dz ? fz : (1.0 - fz);
  const double p = (wX * Y) * Z;
  return (float)((double)acc + p);
}

__device__ __forceinline__ double x_factor(float w, double fx, int dx) {
#pragma clang fp contract(off)
  return (double)w * (dx ? fx : (1.0 - fx));
}

// ---------------------------------------------------------------------------
// Reference-dataflow spread: every cell of the [slot][c][X][Y][Z] grid is
// written (zero outside the ligand box), like memset + FillLigandGrid
// (TranRotaConfSearchNoMini.cu:389-409).  Each cell gathers the channel's
// atoms in ascending order, so the float accumulation order per cell is the
// reference's.
// ---------------------------------------------------------------------------
struct SpreadAtom {
  int ix, iy, iz;
  float w;
  double fx, fy, fz;
};

__global__ __launch_bounds__(kBlock) void k_spread_full(const PoseSlot* __restrict__ slots,
                                                        const LigandDesc* __restrict__ ligs,
                                                        const int* __restrict__ chan_atoms,
                                                        const float* __restrict__ wv, const float* __restrict__ wq,
                                                        const AtomGrid* __restrict__ ag, int astride,
                                                        const RotMeta* __restrict__ meta, int N, int C,
                                                        float* __restrict__ out) {
  __shared__ SpreadAtom sa[kMaxAtoms];
  const int s = blockIdx.z, c = blockIdx.y;
  const size_t G = (size_t)N * N * N;
  float* o = out + ((size_t)s * C + c) * G;
  const size_t p0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const RotMeta m = meta[s];
  const LigandDesc& L = ligs[slots[s].lig];
  const bool active = m.valid && c < L.nch;
  int nc = 0;
  if (active) {
    const int* ca = chan_atoms + L.list_off + L.ch_start[c];
    nc = L.ch_start[c + 1] - L.ch_start[c];
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
      const int a = ca[i];
      const AtomGrid g = ag[(size_t)s * astride + a];
      SpreadAtom t;
      t.ix = g.ix;
      t.iy = g.iy;
      t.iz = g.iz;
      t.w = channel_weight(L, c, a, wv, wq);
      t.fx = g.fx;
      t.fy = g.fy;
      t.fz = g.fz;
      sa[i] = t;
    }
  }
  __syncthreads();
  if (p0 >= G) return;
  const int x = (int)(p0 / ((size_t)N * N));
  const int y = (int)((p0 / N) % N);
  const int z = (int)(p0 % N);
  float acc = 0.0f;
  if (active && x < m.box[0] && y < m.box[1] && z < m.box[2]) {
    for (int i = 0; i < nc; ++i) {
      const int dx = x - sa[i].ix, dy = y - sa[i].iy, dz = z - sa[i].iz;
      if ((unsigned)dx <= 1u && (unsigned)dy <= 1u && (unsigned)dz <= 1u)
        acc = accum_corner(acc, x_factor(sa[i].w, sa[i].fx, dx), sa[i].fy, sa[i].fz, dy, dz);
    }
  }
  o[p0] = acc;
}

// ---------------------------------------------------------------------------
// 3-D transform passes.  Grids are [X][Y][Z] with Z fastest; spectra are
// [X][Y][Z/2+1] (cuFFT's R2C layout, TranRotaConfSearchNoMini.cu:158-174).
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void load_twiddles(float2* tw, const float2* __restrict__ twg) {
  for (int t = threadIdx.x; t < N; t += blockDim.x) tw[t] = twg[t];
}

// Split packed complex rows (row 2p in .x, row 2p+1 in .y, transformed) into
// two Hermitian half spectra rows of length ZH.  Rows >= nrow are zero.
template <int N>
__device__ __forceinline__ void unpack_rows(const float2* __restrict__ r, float2* __restrict__ u, int nrow) {
  constexpr int ZH = N / 2 + 1;
  for (int t = threadIdx.x; t < N * ZH; t += blockDim.x) {
    const int y = t / ZH, k = t - (t / ZH) * ZH;
    float2 v = make_float2(0.f, 0.f);
    if (y < nrow) {
      const int p = y >> 1;
      const float2 zk = r[p * N + k];
      const float2 zm = r[p * N + (k == 0 ? 0 : N - k)];
      if ((y & 1) == 0)
        v = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
      else
        v = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
    }
    u[t] = v;
  }
}

// Pack two Hermitian half-spectrum rows (2p, 2p+1) into one full complex row
// A + iB for a complex inverse transform (C2R two-for-one).  The imaginary
// parts of the DC and Nyquist bins are dropped, as a C2R transform does.
template <int N>
__device__ __forceinline__ void pack_rows_c2r(const float2* __restrict__ r, float2* __restrict__ u, int npair) {
  constexpr int ZH = N / 2 + 1;
  for (int t = threadIdx.x; t < npair * N; t += blockDim.x) {
    const int p = t / N, k = t - (t / N) * N;
    float2 A, B;
    if (k <= N / 2) {
      A = r[(2 * p) * ZH + k];
      B = r[(2 * p + 1) * ZH + k];
      if (k == 0 || k == N / 2) {
        A.y = 0.f;
        B.y = 0.f;
      }
    } else {
      A = r[(2 * p) * ZH + (N - k)];
      B = r[(2 * p + 1) * ZH + (N - k)];
      A.y = -A.y;
      B.y = -B.y;
    }
    u[t] = make_float2(A.x - B.y, A.y + B.x);
  }
}

// Forward plane pass: real plane [Y][Z] -> z R2C of all rows -> y FFT of all
// half-spectrum columns -> [Y][ZH].
template <class P>
__global__ __launch_bounds__(kBlock) void k_r2c_planes(const float* __restrict__ in, float2* __restrict__ out,
                                                       const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  __shared__ float2 sA[N * ZH], sB[N * ZH], tw[N];
  load_twiddles<N>(tw, twg);
  const size_t plane = blockIdx.x;
  const float* ip = in + plane * N * N;
  for (int t = threadIdx.x; t < N * N / 2; t += blockDim.x) {
    const int p = t / N, z = t - (t / N) * N;
    sA[t] = make_float2(ip[(2 * p) * N + z], ip[(2 * p + 1) * N + z]);
  }
  __syncthreads();
  float2* r = lds_fft<P, -1, false>(sA, sB, N / 2, 1, N, tw);
  float2* u = (r == sA) ? sB : sA;
  unpack_rows<N>(r, u, N);
  __syncthreads();
  float2* res = lds_fft<P, -1, true, ZH>(u, r, ZH, ZH, 1, tw);
  float2* op = out + plane * N * ZH;
  for (int t = threadIdx.x; t < N * ZH; t += blockDim.x) op[t] = res[t];
}

// x-line pass over [nb][X][Y][ZH]: one workgroup per (grid, ky) transforms
// the ZH lines of that ky in place.
template <class P, int S>
__global__ __launch_bounds__(kBlock) void k_xlines(float2* __restrict__ data, const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  __shared__ float2 sA[N * ZH], sB[N * ZH], tw[N];
  load_twiddles<N>(tw, twg);
  const size_t b = blockIdx.x / N;
  const int ky = blockIdx.x - (int)b * N;
  float2* base = data + b * N * N * ZH + (size_t)ky * ZH;
  for (int t = threadIdx.x; t < N * ZH; t += blockDim.x) {
    const int x = t / ZH, k = t - (t / ZH) * ZH;
    sA[t] = base[(size_t)x * N * ZH + k];
  }
  __syncthreads();
  float2* res = lds_fft<P, S, true, ZH>(sA, sB, ZH, ZH, 1, tw);
  for (int t = threadIdx.x; t < N * ZH; t += blockDim.x) {
    const int x = t / ZH, k = t - (t / ZH) * ZH;
    base[(size_t)x * N * ZH + k] = res[t];
  }
}

// Inverse plane pass: [Y][ZH] -> y inverse FFT -> z C2R -> real plane [Y][Z].
template <class P>
__global__ __launch_bounds__(kBlock) void k_c2r_planes(const float2* __restrict__ in, float* __restrict__ out,
                                                       const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  __shared__ float2 sA[N * ZH], sB[N * ZH], tw[N];
  load_twiddles<N>(tw, twg);
  const size_t plane = blockIdx.x;
  const float2* ip = in + plane * N * ZH;
  for (int t = threadIdx.x; t < N * ZH; t += blockDim.x) sA[t] = ip[t];
  __syncthreads();
  float2* r = lds_fft<P, 1, true, ZH>(sA, sB, ZH, ZH, 1, tw);
  float2* u = (r == sA) ? sB : sA;
  pack_rows_c2r<N>(r, u, N / 2);
  __syncthreads();
  float2* res = lds_fft<P, 1, false>(u, r, N / 2, 1, N, tw);
  float* op = out + plane * N * N;
  for (int t = threadIdx.x; t < N * N / 2; t += blockDim.x) {
    const int p = t / N, z = t - (t / N) * N;
    const float2 v = res[t];
    op[(2 * p) * N + z] = v.x;
    op[(2 * p + 1) * N + z] = v.y;
  }
}

// ---------------------------------------------------------------------------
// kernel.cu:5-14 ConjMult: L <- conj(L) * P, protein broadcast over poses,
// with the ligand's used-channel map and a bounds guard.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_conj_mult(float2* __restrict__ L, const float2* __restrict__ P,
                                                      const int* __restrict__ pmap, int C, long long H,
                                                      long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long h = i % H;
  const int c = (int)((i / H) % C);
  const float2 l = L[i];
  const float2 p = P[(long long)pmap[c] * H + h];
  L[i] = make_float2(l.x * p.x + l.y * p.y, l.x * p.y - l.y * p.x);
}

// kernel.cu:16-30 SumGrids: S[q] = (sum_c L[q][c]) / sqrtf(G), channels in
// ascending order, float accumulation.
__global__ __launch_bounds__(kBlock) void k_sum_grids(const float2* __restrict__ L, float2* __restrict__ S, int C,
                                                      long long H, long long total, float G) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long q = i / H, h = i - q * H;
  float2 acc = make_float2(0.f, 0.f);
  for (int c = 0; c < C; ++c) {
    const float2 v = L[(q * C + c) * H + h];
    acc.x += v.x;
    acc.y += v.y;
  }
  const float sg = __fsqrt_rn(G);
  S[i] = make_float2(__fdiv_rn(acc.x, sg), __fdiv_rn(acc.y, sg));
}

// TranRotaConfSearchNoMini.cu:441-464 on the device: E = f / sqrt(G) in
// double, stored as float, minimum over the window with first-index ties.
__global__ __launch_bounds__(kBlock) void k_argmin_full(const float* __restrict__ E, const RotMeta* __restrict__ meta,
                                                        unsigned long long* __restrict__ keys, int N) {
  const int s = blockIdx.y, x = blockIdx.x;
  const RotMeta m = meta[s];
  if (!m.valid || x >= m.win[0]) return;
  const size_t G = (size_t)N * N * N;
  const double sg = sqrt((double)G);
  const float* e = E + (size_t)s * G + (size_t)x * N * N;
  const int wy = m.win[1], wz = m.win[2];
  unsigned long long best = ~0ull;
  for (int t = threadIdx.x; t < wy * wz; t += blockDim.x) {
    const int y = t / wz, z = t - (t / wz) * wz;
    const float v = (float)((double)e[y * N + z] / sg);
    const unsigned long long k = energy_key(v, (unsigned)((x * N + y) * N + z));
    best = k < best ? k : best;
  }
  block_min_commit(best, keys + s);
}

// ---------------------------------------------------------------------------
// Fused, pruned pipeline
// ---------------------------------------------------------------------------
struct PlaneAtom {
  int iy, iz;
  double wx, fy, fz;
};

// Line-blocked layout of the protein spectra copy the mid stage reads: lines
// of the (ky, kz) plane in blocks of B = 64/L consecutive lines, the rows x
// of one block contiguous.  Element (x, line) of an R-row volume sits at
// ((line / B) * R + x) * B + line % B, so the 64/L lines a mid-stage
// wavefront owns are 8 x 64/L contiguous bytes per row and its 8 rows of one
// load are 512 contiguous bytes (whole cache lines; the mid stage reads the
// protein spectra once per (pose, channel), twice the ligand's bytes).  The
// ligand planes A and correlation planes T stay plane-major [x][line]: their
// producers and consumers walk planes.  N * (N/2 + 1) is a multiple of B for
// every supported N.
template <int N>
__device__ __forceinline__ int lb_index(int R, int x, int line) {
  constexpr int B = RegShape<N>::LPW;  // a power of two
  return ((line / B) * R + x) * B + (line & (B - 1));  // < 2^31 per (slot, channel)
}

// The protein copy pairs the rows a lane holds next to each other: row
// x = n2 + L*n1 of a block sits at ((n1/2)*L + n2)*2B + 2*line_in_block +
// n1 % 2, so one 16-byte load gives a lane rows n1 = 2p and 2p + 1 of its
// line and one wavefront load covers 1 KB of consecutive bytes.
template <int N>
__device__ __forceinline__ int lbp_index(int x, int line) {
  constexpr int B = RegShape<N>::LPW, L = RegShape<N>::L, V = RegShape<N>::V;
  const int n2 = x % L, n1 = x / L;
  return (((line / B) * (V / 2) + n1 / 2) * L + n2) * (2 * B) + (line & (B - 1)) * 2 + (n1 & 1);
}

template <class P>
__global__ __launch_bounds__(kBlock) void k_to_line_blocks(const float2* __restrict__ in, float2* __restrict__ out,
                                                           int nch) {
  constexpr int N = P::N, NL = N * P::ZH;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t per = (size_t)N * NL;
  if (t >= per * nch) return;
  const int c = (int)(t / per);
  const size_t r = t - (size_t)c * per;
  const int x = (int)(r / NL), line = (int)(r % NL);
  out[(size_t)c * per + (size_t)lbp_index<N>(x, line)] = in[t];
}

// Stage A: one workgroup per (pose, ligand channel).  The channel's atoms are
// staged in LDS once; then for every occupied x-plane: spread the plane in
// LDS (FillLigandGrid arithmetic), z R2C of the occupied rows (two real rows
// per complex line), y FFT of the half-spectrum columns, and write the plane
// spectrum A[slot][c][x][Y][ZH].  No global loads inside the plane loop.
template <class P>
__global__ __launch_bounds__(kBlock) void k_fwd_planes(const PoseSlot* __restrict__ slots,
                                                       const LigandDesc* __restrict__ ligs,
                                                       const int* __restrict__ chan_atoms,
                                                       const float* __restrict__ wv, const float* __restrict__ wq,
                                                       const AtomGrid* __restrict__ ag, int astride,
                                                       const RotMeta* __restrict__ meta, int cmax, int boxmax,
                                                       float2* __restrict__ A, const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  __shared__ float2 sA[N * ZH], sB[N * ZH], tw[N];
  extern __shared__ double dyn[];  // [nmax] SpreadAtom
  const int c = blockIdx.x, s = blockIdx.y;
  const RotMeta m = meta[s];
  if (!m.valid) return;
  const PoseSlot ps = slots[s];
  const LigandDesc& L = ligs[ps.lig];
  if (c >= L.nch) return;
  const int nc = L.ch_start[c + 1] - L.ch_start[c];
  SpreadAtom* sa = reinterpret_cast<SpreadAtom*>(dyn);
  load_twiddles<N>(tw, twg);
  {
    const int* ca = chan_atoms + L.list_off + L.ch_start[c];
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
      const int a = ca[i];
      const AtomGrid g = ag[(size_t)s * astride + a];
      SpreadAtom t;
      t.ix = g.ix;
      t.iy = g.iy;
      t.iz = g.iz;
      t.w = channel_weight(L, c, a, wv, wq);
      t.fx = g.fx;
      t.fy = g.fy;
      t.fz = g.fz;
      sa[i] = t;
    }
  }
  const int bxm = min(m.box[0], boxmax), by = m.box[1];
  const int npair = (by + 1) >> 1;
  float2* op0 = A + (((size_t)s * cmax + c) * boxmax) * (size_t)(N * ZH);
  for (int x = 0; x < bxm; ++x) {
    __syncthreads();  // atoms staged; previous plane written out
    // Zero the packed rows (row pair p holds rows 2p in .x, 2p+1 in .y).
    for (int t = threadIdx.x; t < npair * N; t += blockDim.x) sA[t] = make_float2(0.f, 0.f);
    __syncthreads();
    // Ordered scatter by one wavefront: atoms touching plane x in ascending
    // order, the 4 corners of one atom in 4 lanes (distinct cells), so every
    // cell accumulates its contributions in the reference's atom order
    // (FillLigandGrid.cpp:45-66, 97-117).
    if (threadIdx.x < 64) {
      const int lane = threadIdx.x;
      const int dy = (lane >> 1) & 1, dz = lane & 1;
      float* cells = reinterpret_cast<float*>(sA);
      for (int base = 0; base < nc; base += 64) {
        const int aix = base + lane < nc ? sa[base + lane].ix : -4;
        unsigned long long mk = __ballot(x - aix == 0 || x - aix == 1);
        while (mk) {
          const int i = base + __ffsll((long long)mk) - 1;
          mk &= mk - 1;
          const int dx = x - sa[i].ix;
          if (lane < 4) {
            const int y = sa[i].iy + dy, z = sa[i].iz + dz;
            float* cell = cells + (((y >> 1) * N + z) << 1) + (y & 1);
            *cell = accum_corner(*cell, x_factor(sa[i].w, sa[i].fx, dx), sa[i].fy, sa[i].fz, dy, dz);
          }
          WaveSync::sync();
        }
      }
    }
    __syncthreads();
    float2* r = lds_fft<P, -1, false>(sA, sB, npair, 1, N, tw);
    float2* u = (r == sA) ? sB : sA;
    unpack_rows<N>(r, u, 2 * npair);
    __syncthreads();
    const float2* res = lds_fft<P, -1, true, ZH>(u, r, ZH, ZH, 1, tw);
    for (int t = threadIdx.x; t < N * ZH; t += blockDim.x) op0[(size_t)x * (N * ZH) + t] = res[t];
  }
}

// Stage A, register form: one workgroup per (pose, ligand channel) stages
// the channel's atoms once, then per occupied x-plane
//   1. zeroes the packed row pairs of an LDS tile (pair p = rows 2p, 2p+1 in
//      .x, .y) and spreads the plane's cells into them in the reference's
//      atom order (one wavefront, atoms selected by ballot),
//   2. z-transforms the row pairs with register four-step FFTs (the four
//      wavefronts take 64/L pairs each; first stage pruned when the box fits
//      in N/2) and writes each packed spectrum back over its row,
//   3. y-transforms the half-spectrum columns, unpacking the two real rows of
//      every pair as it loads them (columns 0 and N/2 are real along y and
//      share one complex line), and writes the plane spectrum to A.
// Four block barriers per plane; the tile holds only the box's row pairs
// (dynamic LDS), so every box size fits.  Tile rows have stride
// RSTR = N + pad with RSTR % 32 == 4 (float2 units): the z-stage row loads of
// 32 consecutive lanes then touch 32 distinct 8-byte bank slots.
template <int N>
struct FwdRegShape {
  static constexpr int RSTR = N + ((4 - N % 32) + 32) % 32;
};

// One x-plane of the register forward stages (steps 1 and 2 above): zero the
// tile's row pairs, spread the plane's cells into them in the reference's
// atom order (wavefront 0, atoms touching the plane picked by ballot;
// FillLigandGrid.cpp:45-66, 97-117), z-transform the row pairs in place.
// Ends behind a block barrier with the packed row spectra in the tile.
template <int N>
__device__ __forceinline__ void fwd_plane_rows(int x, const SpreadAtom* __restrict__ sa, int nc, float2* zt,
                                               int npair, bool zhalf, const float2* twl, float2* tr, int wave,
                                               int lane) {
  using RS_ = RegShape<N>;
  constexpr int L = RS_::L, V = RS_::V, LPW = RS_::LPW, RSTR = FwdRegShape<N>::RSTR;
  constexpr bool PRUNE = V == 8 || V == 16;
  const int ln = lane % LPW, n2 = lane / LPW;
  const int zsh = zhalf ? 31 - __builtin_clz(N / 2) : 31 - __builtin_clz(N);  // log2 of the row length zeroed
  const int dy = (lane >> 1) & 1, dz = lane & 1;
  static_assert((N & (N - 1)) == 0 || !PRUNE, "pruned lengths are powers of two");
  __syncthreads();  // atoms staged; the previous plane's reads of the tile are done
  if ((N & (N - 1)) == 0) {
    for (int t = threadIdx.x; t < (npair << zsh); t += blockDim.x) {
      const int p = t >> zsh;
      zt[p * RSTR + (t - (p << zsh))] = make_float2(0.f, 0.f);
    }
  } else {
    for (int t = threadIdx.x; t < npair * N; t += blockDim.x) {
      const int p = t / N;
      zt[p * RSTR + (t - p * N)] = make_float2(0.f, 0.f);
    }
  }
  __syncthreads();
  if (wave == 0) {
    float* cells = reinterpret_cast<float*>(zt);
    for (int base = 0; base < nc; base += 64) {
      const int aix = base + lane < nc ? sa[base + lane].ix : -4;
      unsigned long long mk = __ballot(x - aix == 0 || x - aix == 1);
      while (mk) {
        const int i = base + __ffsll((long long)mk) - 1;
        mk &= mk - 1;
        const int ddx = x - sa[i].ix;
        if (lane < 4) {
          const int y = sa[i].iy + dy, z = sa[i].iz + dz;
          float* cell = cells + (((y >> 1) * RSTR + z) << 1) + (y & 1);
          *cell = accum_corner(*cell, x_factor(sa[i].w, sa[i].fx, ddx), sa[i].fy, sa[i].fz, dy, dz);
        }
        WaveSync::sync();
      }
    }
  }
  __syncthreads();
  for (int p0 = wave * LPW; p0 < npair; p0 += 4 * LPW) {
    const int p = p0 + ln;
    const bool ok = p < npair;
    float2 v[V];
    if (PRUNE && zhalf) {
#pragma unroll
      for (int n1 = 0; n1 < V / 2; ++n1) v[n1] = ok ? zt[p * RSTR + n2 + L * n1] : make_float2(0.f, 0.f);
      reg_fft<N, -1, PRUNE>(v, twl, tr, ln, n2);
    } else {
#pragma unroll
      for (int n1 = 0; n1 < V; ++n1) v[n1] = ok ? zt[p * RSTR + n2 + L * n1] : make_float2(0.f, 0.f);
      reg_fft<N, -1, false>(v, twl, tr, ln, n2);
    }
    if (ok) {
#pragma unroll
      for (int q = 0; q < V; ++q) zt[p * RSTR + n2 + L * q] = v[q];
    }
  }
  __syncthreads();
}

template <class P>
__global__ __launch_bounds__(kBlock) void k_fwd_planes_reg(const PoseSlot* __restrict__ slots,
                                                           const LigandDesc* __restrict__ ligs,
                                                           const int* __restrict__ chan_atoms,
                                                           const float* __restrict__ wv, const float* __restrict__ wq,
                                                           const AtomGrid* __restrict__ ag, int astride,
                                                           const RotMeta* __restrict__ meta, int cmax, int boxmax,
                                                           int natom_lds, int npmax, float2* __restrict__ A,
                                                           const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  using RS_ = RegShape<N>;
  constexpr int L = RS_::L, V = RS_::V, LPW = RS_::LPW, RSTR = FwdRegShape<N>::RSTR;
  constexpr bool TW_LDS = V >= 16;  // long lines: twiddles from LDS, registers to the data
  __shared__ float2 sbuf[4][RS_::SLICE];
  __shared__ float2 stw[TW_LDS ? L * V : 1];
  extern __shared__ double dyn[];  // [natom_lds] SpreadAtom, then the tile [npmax][RSTR]
  const int c = blockIdx.x, s = blockIdx.y;
  const RotMeta m = meta[s];
  if (!m.valid) return;
  const PoseSlot ps = slots[s];
  const LigandDesc& Ld = ligs[ps.lig];
  if (c >= Ld.nch) return;
  const int nc = Ld.ch_start[c + 1] - Ld.ch_start[c];
  SpreadAtom* sa = reinterpret_cast<SpreadAtom*>(dyn);
  float2* zt = reinterpret_cast<float2*>(sa + natom_lds);
  {
    const int* ca = chan_atoms + Ld.list_off + Ld.ch_start[c];
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
      const int a = ca[i];
      const AtomGrid g = ag[(size_t)s * astride + a];
      SpreadAtom t;
      t.ix = g.ix;
      t.iy = g.iy;
      t.iz = g.iz;
      t.w = channel_weight(Ld, c, a, wv, wq);
      t.fx = g.fx;
      t.fy = g.fy;
      t.fz = g.fz;
      sa[i] = t;
    }
  }
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int ln = lane % LPW, n2 = lane / LPW;
  float2* tr = sbuf[wave];
  float2 twr[TW_LDS ? 1 : V];
  if constexpr (TW_LDS) {
    for (int t = threadIdx.x; t < L * V; t += blockDim.x) stw[t] = twg[(t / V) * (t % V)];
  } else {
    reg_twiddles<N>(twg, n2, twr);
  }
  const float2* twl = TW_LDS ? stw + n2 * V : twr;
  const int bxm = min(m.box[0], boxmax), by = m.box[1], bz = m.box[2];
  const int npair = (by + 1) >> 1;
  // A pruned first stage exists for 8 and 16 values per lane; other lengths
  // take the full transform (one code path, fewer live registers).
  constexpr bool PRUNE = V == 8 || V == 16;
  const bool zhalf = PRUNE && bz <= N / 2, yhalf = PRUNE && by <= N / 2;
  float2* op0 = A + (((size_t)s * cmax + c) * boxmax) * (size_t)(N * ZH);
  for (int x = 0; x < bxm; ++x) {
    fwd_plane_rows<N>(x, sa, nc, zt, npair, zhalf, twl, tr, wave, lane);
    // 3. y transforms; line j = column j, except line 0 = columns 0 (real
    // part) and N/2 (imaginary part).
    for (int j0 = wave * LPW; j0 < N / 2; j0 += 4 * LPW) {
      const int j = j0 + ln;
      const bool jok = j < N / 2;
      const int ia = j == 0 ? 0 : j, ib = j == 0 ? N / 2 : N - j;
      float2 v[V];
      auto col = [&](int y) {
        const int p = y >> 1;
        float2 r = make_float2(0.f, 0.f);
        if (p < npair && jok) {
          const float2 zk = zt[p * RSTR + ia], zm = zt[p * RSTR + ib];
          if (j == 0)
            r = (y & 1) ? make_float2(zk.y, zm.y) : make_float2(zk.x, zm.x);
          else
            r = (y & 1) ? make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x))
                        : make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        }
        return r;
      };
      if (PRUNE && yhalf) {
#pragma unroll
        for (int n1 = 0; n1 < V / 2; ++n1) v[n1] = col(n2 + L * n1);
        reg_fft<N, -1, PRUNE>(v, twl, tr, ln, n2);
      } else {
#pragma unroll
        for (int n1 = 0; n1 < V; ++n1) v[n1] = col(n2 + L * n1);
        reg_fft<N, -1, false>(v, twl, tr, ln, n2);
      }
      // v[q] = Y[ky = n2 + L q] of line j.
      if (j0 == 0) {
        // Split line 0: C0(ky) = (U(ky) + conj U(-ky)) / 2,
        //               CN(ky) = (U(ky) - conj U(-ky)) / 2i.
        WaveSync::sync();
        if (j == 0) {
#pragma unroll
          for (int q = 0; q < V; ++q) tr[n2 + L * q] = v[q];
        }
        WaveSync::sync();
        if (j == 0) {
#pragma unroll
          for (int q = 0; q < V; ++q) {
            const int ky = n2 + L * q;
            const float2 u = v[q], w = tr[ky == 0 ? 0 : N - ky];
            op0[(size_t)x * (N * ZH) + ky * ZH] = make_float2(0.5f * (u.x + w.x), 0.5f * (u.y - w.y));
            op0[(size_t)x * (N * ZH) + ky * ZH + N / 2] = make_float2(0.5f * (u.y + w.y), -0.5f * (u.x - w.x));
          }
        }
      }
      if (j > 0 && jok) {
#pragma unroll
        for (int q = 0; q < V; ++q) op0[(size_t)x * (N * ZH) + (n2 + L * q) * ZH + j] = v[q];
      }
    }
  }
}

// Stage B: x FFT of every used channel, conj-multiply with the protein
// spectra, channel sum, 1/G, inverse x FFT, for the (ky, kz) lines of the
// spectrum.  Each wavefront owns 64/L consecutive lines (flattened
// ky*ZH + kz) held in registers, L lanes per line (register four-step FFT,
// one swizzled LDS transpose per transform).  A line's data arrives from
// global memory already in the register layout the transform wants, and the
// forward output layout is the inverse's input layout.  A workgroup is 4
// wavefronts and walks a group of poses as one sequence of (pose, channel)
// items; a ring of DEPTH register buffers (DEPTH tuned, 1 for long lines)
// keeps the next items' loads in flight while one item is transformed.  When the
// pose's box is at most N/2 the upper half of every line is zero: those rows
// are not loaded and the first DFT stage is pruned.  Workgroups map to
// (line group, pose group) by xcd_unit_of, so each XCD touches about one
// eighth of the protein spectra, which then stays in that XCD's L2.  Reads
// A[slot][c][x][line] for x < box, writes T[slot][x'][line] for x' < window.

#ifdef CDK_MID_WAVES
#define CDK_MID_ATTR __attribute__((amdgpu_waves_per_eu(CDK_MID_WAVES)))
#else
#define CDK_MID_ATTR
#endif
template <class P, int DEPTH, bool BOXHALF>
__global__ __launch_bounds__(kBlock) CDK_MID_ATTR void k_mid(const PoseSlot* __restrict__ slots, const LigandDesc* __restrict__ ligs,
                                                const RotMeta* __restrict__ meta, int nslots, int group,
                                                const float2* __restrict__ A, int cmax, int boxmax,
                                                const float2* __restrict__ Phat, float2* __restrict__ T, float invG,
                                                const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  using RS = RegShape<N>;
  constexpr int L = RS::L, V = RS::V, LPW = RS::LPW;
  constexpr int NL = N * ZH;             // lines in a spectrum
  constexpr int LPB = 4 * LPW;           // lines per workgroup
  constexpr int NLG = (NL + LPB - 1) / LPB;
  // Long lines read their twiddles from LDS (registers go to the data).
  constexpr bool TW_LDS = V >= 16;
  // BOXHALF: every pose's box fits in N/2, so only V/2 ligand rows exist.
  constexpr int VA = BOXHALF ? V / 2 : V;
  __shared__ float2 sbuf[4][RS::SLICE];
  __shared__ float2 stw[TW_LDS ? L * V : 1];
  int lg, sg;
  if (!xcd_unit_of(blockIdx.x, NLG / 8, NLG % 8, (nslots + group - 1) / group, lg, sg)) return;
  if constexpr (TW_LDS) {
    for (int t = threadIdx.x; t < L * V; t += blockDim.x) stw[t] = twg[(t / V) * (t % V)];
    __syncthreads();
  }
  // Wave index and line offset are wave-uniform; say so, so the buffer
  // descriptors built from them live in scalar registers.
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int ln = lane % LPW, n2 = lane / LPW;
  const int l0 = __builtin_amdgcn_readfirstlane(lg * LPB + wave * LPW);
  if (l0 >= NL) return;
  const bool lok = l0 + ln < NL;
  const size_t plane = (size_t)NL;
  float2 twr[TW_LDS ? 1 : V];
  if constexpr (!TW_LDS) reg_twiddles<N>(twg, n2, twr);
  const float2* twl = TW_LDS ? stw + n2 * V : twr;
  float2* lds = sbuf[wave];
  const int s1 = min(nslots, sg * group + group);
  // Ligand planes A and correlation planes T are plane-major: lane offset of
  // row x = n2 of the wavefront's lines, row n1 adds L*n1 planes.  The
  // protein copy is line-blocked with paired rows (lbp_index): block
  // l0 / LPW, 16 bytes = rows n1 = 2p, 2p+1 per lane.  Rows past a
  // descriptor's end (x >= box, x' >= window) read as zero / are dropped.
  const size_t blk = (size_t)(l0 / LPW) * LPW;  // protein block start, in rows of LPW lines
  const int voffA = lok ? (n2 * NL + ln) * (int)sizeof(float2) : kVoffNone;
  constexpr int kRowA = L * NL * (int)sizeof(float2);
  const int voffP = lok ? (n2 * 2 * LPW + ln * 2) * (int)sizeof(float2) : kVoffNone;
  constexpr int kPairP = L * 2 * LPW * (int)sizeof(float2);  // rows n1 = 2p, 2p+1 -> pair p
  // Slot-derived values are wave-uniform; readfirstlane keeps them (and the
  // descriptors built from them) in scalar registers.
  auto uni = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
  auto next_valid = [&](int s) {
    while (s < s1 && !uni(meta[s].valid)) ++s;
    return s;
  };
  // One (pose, channel) item of the sequence.
  struct Item {
    int s, c, nch, bx, lig;
  };
  auto enter = [&](Item& it, int s) {
    it.s = s;
    it.c = 0;
    if (s < s1) {
      it.lig = uni(slots[s].lig);
      it.nch = uni(ligs[it.lig].nch);
      it.bx = uni(min(meta[s].box[0], boxmax));
    }
  };
  auto advance = [&](Item& it) {
    if (++it.c >= it.nch) enter(it, next_valid(it.s + 1));
  };
  auto load = [&](const Item& it, float2* ra, float2* rp) {
    const __amdgpu_buffer_rsrc_t r =
        buffer_rsrc(A + (((size_t)it.s * cmax + it.c) * boxmax) * plane + l0, (unsigned)(it.bx * NL * sizeof(float2)));
    const int pc = uni(ligs[it.lig].pmap[it.c]);
    const __amdgpu_buffer_rsrc_t q =
        buffer_rsrc(Phat + (size_t)pc * N * plane + blk * N, (unsigned)(N * LPW * sizeof(float2)));
    if (BOXHALF || it.bx <= N / 2) {
#pragma unroll
      for (int n1 = 0; n1 < V / 2; ++n1) ra[n1] = buffer_load_f2(r, voffA + n1 * kRowA);
    } else {
#pragma unroll
      for (int n1 = 0; n1 < VA; ++n1) ra[n1] = buffer_load_f2(r, voffA + n1 * kRowA);
    }
#pragma unroll
    for (int m = 0; m < V / 2; ++m) {
      const float4 f = buffer_load_f4(q, voffP + m * kPairP);
      rp[2 * m] = make_float2(f.x, f.y);
      rp[2 * m + 1] = make_float2(f.z, f.w);
    }
  };
  float2 acc[V];
#pragma unroll
  for (int i = 0; i < V; ++i) acc[i] = make_float2(0.f, 0.f);
  // Transform item `it` held in (ra, rp) and accumulate conj(L) P; after the
  // pose's last channel run the inverse and store the window rows.
  auto consume = [&](const Item& it, float2* ra, const float2* rp) {
    // The transform runs in place on a full buffer; a half buffer (BOXHALF)
    // widens into a temporary.
    float2 vt[VA < V ? V : 1];
    float2* v = ra;
    if constexpr (VA < V) {
#pragma unroll
      for (int i = 0; i < VA; ++i) vt[i] = ra[i];
      v = vt;
    }
    if (BOXHALF || it.bx <= N / 2)
      reg_fft<N, -1, true>(v, twl, lds, ln, n2);
    else
      reg_fft<N, -1, false>(v, twl, lds, ln, n2);
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const float2 l = v[i], p = rp[i];
      acc[i].x = fmaf(l.x, p.x, acc[i].x);
      acc[i].x = fmaf(l.y, p.y, acc[i].x);
      acc[i].y = fmaf(l.x, p.y, acc[i].y);
      acc[i].y = fmaf(-l.y, p.x, acc[i].y);
    }
    if (it.c + 1 == it.nch) {
#pragma unroll
      for (int i = 0; i < V; ++i) acc[i] = c_scale(acc[i], invG);
      reg_fft<N, 1>(acc, twl, lds, ln, n2);
      const int wx = uni(meta[it.s].win[0]);
      const __amdgpu_buffer_rsrc_t rt =
          buffer_rsrc(T + (size_t)it.s * N * plane + l0, (unsigned)(max(wx, 0) * NL * sizeof(float2)));
#pragma unroll
      for (int m = 0; m < V; ++m) buffer_store_f2(acc[m], rt, voffA + m * kRowA);
#pragma unroll
      for (int i = 0; i < V; ++i) acc[i] = make_float2(0.f, 0.f);
    }
  };
  Item first;
  enter(first, next_valid(sg * group));
  if (first.s >= s1) return;
  // Ring of DEPTH register buffers over the (pose, channel) sequence: the
  // loads of the next DEPTH-1 items are in flight while one is transformed.
  // Buffer indices are compile-time (the ring step is unrolled), so nothing
  // is copied between registers.
  Item it[DEPTH];
  float2 ab[DEPTH][VA], pb[DEPTH][V];
  it[0] = first;
#pragma unroll
  for (int d = 1; d < DEPTH; ++d) {
    it[d] = it[d - 1];
    advance(it[d]);
  }
#pragma unroll
  for (int d = 0; d < DEPTH; ++d)
    if (it[d].s < s1) load(it[d], ab[d], pb[d]);
  for (;;) {
#pragma unroll
    for (int d = 0; d < DEPTH; ++d) {
      if (it[d].s >= s1) return;  // items are in order: nothing after it is left
      consume(it[d], ab[d], pb[d]);
      Item nx = it[(d + DEPTH - 1) % DEPTH];
      advance(nx);
      it[d] = nx;
      if (nx.s < s1) load(nx, ab[d], pb[d]);
    }
  }
}

// Stage B with the protein spectra shared by several poses: the wavefront
// takes a run of up to NP consecutive valid slots of one ligand (same
// channels, same protein channel map) and walks channels in the outer loop,
// so one protein load serves NP poses.  Each pose keeps its own accumulator
// and sums its channels in the same order with the same operations as
// k_mid, so the two kernels give bit-identical results.  A ring of two
// steps (a step = one channel of one run: the protein rows plus NP ligand
// row sets) keeps the next step's loads in flight during the current
// step's NP transforms.  Lines of at most 8 values per lane only.
template <class P, int NP, bool BOXHALF>
__global__ __launch_bounds__(kBlock) CDK_MID_ATTR void k_mid_np(const PoseSlot* __restrict__ slots,
                                                   const LigandDesc* __restrict__ ligs,
                                                   const RotMeta* __restrict__ meta, int nslots, int group,
                                                   const float2* __restrict__ A, int cmax, int boxmax,
                                                   const float2* __restrict__ Phat, float2* __restrict__ T,
                                                   float invG, const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  using RS = RegShape<N>;
  constexpr int L = RS::L, V = RS::V, LPW = RS::LPW;
  static_assert(V <= 8, "k_mid_np keeps NP accumulators in registers: short lines only");
  constexpr int NL = N * ZH;
  constexpr int LPB = 4 * LPW;
  constexpr int NLG = (NL + LPB - 1) / LPB;
  constexpr int VA = BOXHALF ? V / 2 : V;
  __shared__ float2 sbuf[4][RS::SLICE];
  int lg, sg;
  if (!xcd_unit_of(blockIdx.x, NLG / 8, NLG % 8, (nslots + group - 1) / group, lg, sg)) return;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int ln = lane % LPW, n2 = lane / LPW;
  const int l0 = __builtin_amdgcn_readfirstlane(lg * LPB + wave * LPW);
  if (l0 >= NL) return;
  const bool lok = l0 + ln < NL;
  const size_t plane = (size_t)NL;
  float2 twr[V];
  reg_twiddles<N>(twg, n2, twr);
  float2* lds = sbuf[wave];
  const int s1 = min(nslots, sg * group + group);
  const size_t blk = (size_t)(l0 / LPW) * LPW;
  const int voffA = lok ? (n2 * NL + ln) * (int)sizeof(float2) : kVoffNone;
  constexpr int kRowA = L * NL * (int)sizeof(float2);
  const int voffP = lok ? (n2 * 2 * LPW + ln * 2) * (int)sizeof(float2) : kVoffNone;
  constexpr int kPairP = L * 2 * LPW * (int)sizeof(float2);
  auto uni = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
  auto next_valid = [&](int s) {
    while (s < s1 && !uni(meta[s].valid)) ++s;
    return s;
  };
  // Channel c of a run of np <= NP slots s[0..np) of ligand lig; np == 0 ends
  // the sequence.  next = the first slot after the run.
  struct Step {
    int s[NP], bx[NP];
    int np, c, nch, lig, next;
  };
  auto run_from = [&](Step& st, int s) {
    st.c = 0;
    st.np = 0;
    s = next_valid(s);
    st.next = s;
    if (s >= s1) return;
    st.lig = uni(slots[s].lig);
    st.nch = uni(ligs[st.lig].nch);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      st.s[p] = s;
      st.bx[p] = 0;
      if (st.np == p && s < s1 && uni(slots[s].lig) == st.lig) {
        st.bx[p] = uni(min(meta[s].box[0], boxmax));
        st.np = p + 1;
        s = next_valid(s + 1);
      }
    }
    st.next = s;
  };
  auto advance = [&](Step& st) {
    if (st.np > 0 && ++st.c >= st.nch) run_from(st, st.next);
  };
  auto load = [&](const Step& st, float2 (*ra)[VA], float2* rp) {
    const int pc = uni(ligs[st.lig].pmap[st.c]);
    const __amdgpu_buffer_rsrc_t q =
        buffer_rsrc(Phat + (size_t)pc * N * plane + blk * N, (unsigned)(N * LPW * sizeof(float2)));
#pragma unroll
    for (int m = 0; m < V / 2; ++m) {
      const float4 f = buffer_load_f4(q, voffP + m * kPairP);
      rp[2 * m] = make_float2(f.x, f.y);
      rp[2 * m + 1] = make_float2(f.z, f.w);
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      if (p < st.np) {
        const __amdgpu_buffer_rsrc_t r =
            buffer_rsrc(A + (((size_t)st.s[p] * cmax + st.c) * boxmax) * plane + l0,
                        (unsigned)(st.bx[p] * NL * sizeof(float2)));
        if (BOXHALF || st.bx[p] <= N / 2) {
#pragma unroll
          for (int n1 = 0; n1 < V / 2; ++n1) ra[p][n1] = buffer_load_f2(r, voffA + n1 * kRowA);
        } else {
#pragma unroll
          for (int n1 = 0; n1 < VA; ++n1) ra[p][n1] = buffer_load_f2(r, voffA + n1 * kRowA);
        }
      }
    }
  };
  float2 acc[NP][V];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int i = 0; i < V; ++i) acc[p][i] = make_float2(0.f, 0.f);
  auto consume = [&](const Step& st, float2 (*ra)[VA], const float2* rp) {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      if (p < st.np) {
        float2 v[V];
#pragma unroll
        for (int i = 0; i < VA; ++i) v[i] = ra[p][i];
        if (BOXHALF || st.bx[p] <= N / 2)
          reg_fft<N, -1, true>(v, twr, lds, ln, n2);
        else
          reg_fft<N, -1, false>(v, twr, lds, ln, n2);
#pragma unroll
        for (int i = 0; i < V; ++i) {
          const float2 l = v[i], q = rp[i];
          acc[p][i].x = fmaf(l.x, q.x, acc[p][i].x);
          acc[p][i].x = fmaf(l.y, q.y, acc[p][i].x);
          acc[p][i].y = fmaf(l.x, q.y, acc[p][i].y);
          acc[p][i].y = fmaf(-l.y, q.x, acc[p][i].y);
        }
      }
    }
    if (st.c + 1 == st.nch) {
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        if (p < st.np) {
#pragma unroll
          for (int i = 0; i < V; ++i) acc[p][i] = c_scale(acc[p][i], invG);
          reg_fft<N, 1>(acc[p], twr, lds, ln, n2);
          const int wx = uni(meta[st.s[p]].win[0]);
          const __amdgpu_buffer_rsrc_t rt =
              buffer_rsrc(T + (size_t)st.s[p] * N * plane + l0, (unsigned)(max(wx, 0) * NL * sizeof(float2)));
#pragma unroll
          for (int m = 0; m < V; ++m) buffer_store_f2(acc[p][m], rt, voffA + m * kRowA);
        }
#pragma unroll
        for (int i = 0; i < V; ++i) acc[p][i] = make_float2(0.f, 0.f);
      }
    }
  };
  Step st[2];
  run_from(st[0], sg * group);
  if (st[0].np == 0) return;
  st[1] = st[0];
  advance(st[1]);
  float2 ab[2][NP][VA], pb[2][V];
  load(st[0], ab[0], pb[0]);
  if (st[1].np > 0) load(st[1], ab[1], pb[1]);
  for (;;) {
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      if (st[d].np == 0) return;  // steps are in order: nothing after it is left
      consume(st[d], ab[d], pb[d]);
      Step nx = st[d ^ 1];
      advance(nx);
      st[d] = nx;
      if (nx.np > 0) load(nx, ab[d], pb[d]);
    }
  }
}

// Stage C: one workgroup per (pose, window x').  Inverse y FFT of the plane,
// z C2R of the window rows only, argmin over the (y', z') window.
template <class P>
__global__ __launch_bounds__(kBlock) void k_inv_argmin(const RotMeta* __restrict__ meta,
                                                       const float2* __restrict__ T,
                                                       unsigned long long* __restrict__ keys,
                                                       const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  __shared__ float2 sA[N * ZH], sB[N * ZH], tw[N];
  const int xp = blockIdx.x, s = blockIdx.y;
  const RotMeta m = meta[s];
  if (!m.valid || xp >= m.win[0]) return;
  load_twiddles<N>(tw, twg);
  const float2* ip = T + ((size_t)s * N + xp) * (size_t)(N * ZH);
  for (int t = threadIdx.x; t < N * ZH; t += blockDim.x) sA[t] = ip[t];
  __syncthreads();
  float2* r = lds_fft<P, 1, true, ZH>(sA, sB, ZH, ZH, 1, tw);
  float2* u = (r == sA) ? sB : sA;
  const int wy = m.win[1], wz = m.win[2];
  const int npair = (wy + 1) >> 1;
  pack_rows_c2r<N>(r, u, npair);
  __syncthreads();
  const float2* res = lds_fft<P, 1, false>(u, r, npair, 1, N, tw);
  unsigned long long best = ~0ull;
  for (int t = threadIdx.x; t < wy * wz; t += blockDim.x) {
    const int y = t / wz, z = t - (t / wz) * wz;
    const float2 v2 = res[(y >> 1) * N + z];
    const float v = (y & 1) ? v2.y : v2.x;
    const unsigned long long k = energy_key(v, (unsigned)((xp * N + y) * N + z));
    best = k < best ? k : best;
  }
  block_min_commit(best, keys + s);
}

// Stage C, register form: one workgroup per (pose, group of XPB window
// planes x').  Per plane the four wavefronts
//   1. inverse-transform the half-spectrum columns along y (register
//      four-step FFT; columns 0 and N/2 are Hermitian along y, so their
//      transforms are real and share one complex line),
//   2. write rows y < window into an LDS tile [y][kz] (one block barrier),
//   3. C2R the row pairs along z (two real rows per complex line, the
//      Hermitian half extended on the fly),
// keep a running (energy, index) key minimum over the window, and commit it
// once per workgroup with atomicMin.  The next plane's column loads are
// issued before step 3.  Tile rows have stride RS = ZH + pad with RS % 32 =
// 4 (8 lanes per line) or 2 (4 lanes per line) float2, so both the row
// writes of step 2 and the paired row reads of step 3 spread over the
// 8-byte bank slots.
template <int N>
struct InvRegShape {
  static constexpr int L = RegShape<N>::L;
  static constexpr int ZH = N / 2 + 1;
  static constexpr int T32 = L == 8 ? 4 : 2;
  static constexpr int RS = ZH + ((T32 - ZH % 32) + 32) % 32;
  static constexpr int XPB = 4;  // window planes per workgroup
};

template <class P>
__global__ __launch_bounds__(kBlock) void k_inv_argmin_reg(const RotMeta* __restrict__ meta,
                                                           const float2* __restrict__ T,
                                                           unsigned long long* __restrict__ keys,
                                                           const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  using RS_ = RegShape<N>;
  using IS = InvRegShape<N>;
  constexpr int L = RS_::L, V = RS_::V, LPW = RS_::LPW, RS = IS::RS, XPB = IS::XPB;
  constexpr int NYL = N / 2;  // y lines: columns 1..N/2-1 and the packed {0, N/2}
  // Long lines (>= 16 values per lane): twiddles from LDS and no prefetch of
  // the next plane, so the registers go to the data.
  constexpr bool LONG = V >= 16;
  __shared__ float2 tile[N * RS];
  __shared__ float2 sbuf[4][RS_::SLICE];
  __shared__ float2 stw[LONG ? L * V : 1];
  const int s = blockIdx.y;
  const RotMeta m = meta[s];
  const int x0 = blockIdx.x * XPB;
  if (!m.valid || x0 >= m.win[0]) return;
  if constexpr (LONG) {
    for (int t = threadIdx.x; t < L * V; t += blockDim.x) stw[t] = twg[(t / V) * (t % V)];
    __syncthreads();
  }
  const int x1 = min(x0 + XPB, m.win[0]);
  const int wy = m.win[1], wz = m.win[2];
  const int npw = (wy + 1) >> 1;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int ln = lane % LPW, n2 = lane / LPW;
  float2 twr[LONG ? 1 : V];
  if constexpr (!LONG) reg_twiddles<N>(twg, n2, twr);
  const float2* twl = LONG ? stw + n2 * V : twr;
  float2* lds = sbuf[wave];
  unsigned long long best = ~0ull;
  auto load_cols = [&](int xp, int pass, float2* v) {
    const int j = (pass * 4 + wave) * LPW + ln;
    const float2* ip = T + ((size_t)s * N + xp) * (size_t)(N * ZH);
#pragma unroll
    for (int n1 = 0; n1 < V; ++n1) {
      const int ky = n2 + L * n1;
      float2 r = make_float2(0.f, 0.f);
      if (j == 0) {
        const float2 a = ip[ky * ZH], b = ip[ky * ZH + N / 2];
        r = make_float2(a.x - b.y, a.y + b.x);
      } else if (j < NYL) {
        r = ip[ky * ZH + j];
      }
      v[n1] = r;
    }
  };
  constexpr int YPASS = (NYL + 4 * LPW - 1) / (4 * LPW);
  float2 nxt[LONG ? 1 : V];
  if constexpr (!LONG) load_cols(x0, 0, nxt);
  for (int xp = x0; xp < x1; ++xp) {
    // 1-2. y inverse transforms into the tile.
#pragma unroll
    for (int pass = 0; pass < YPASS; ++pass) {
      float2 v[V];
      if (!LONG && pass == 0) {
#pragma unroll
        for (int i = 0; i < (LONG ? 1 : V); ++i) v[i] = nxt[i];
      } else {
        load_cols(xp, pass, v);
      }
      reg_fft<N, 1>(v, twl, lds, ln, n2);
      const int j = (pass * 4 + wave) * LPW + ln;
      if (j < NYL) {
#pragma unroll
        for (int q = 0; q < V; ++q) {
          const int y = n2 + L * q;
          if (y < 2 * npw) {
            if (j == 0) {
              tile[y * RS] = make_float2(v[q].x, 0.f);
              tile[y * RS + N / 2] = make_float2(v[q].y, 0.f);
            } else {
              tile[y * RS + j] = v[q];
            }
          }
        }
      }
    }
    __syncthreads();
    if constexpr (!LONG) {
      if (xp + 1 < x1) load_cols(xp + 1, 0, nxt);
    }
    // 3. z C2R of the row pairs and the window minimum.
    for (int p0 = wave * LPW; p0 < npw; p0 += 4 * LPW) {
      const int p = p0 + ln;
      const bool ok = p < npw;
      float2 z[V];
#pragma unroll
      for (int n1 = 0; n1 < V; ++n1) {
        const int k = n2 + L * n1;
        float2 r = make_float2(0.f, 0.f);
        if (ok) {
          const int kk = k <= N / 2 ? k : N - k;
          float2 A = tile[(2 * p) * RS + kk], B = tile[(2 * p + 1) * RS + kk];
          if (k > N / 2) {
            A.y = -A.y;
            B.y = -B.y;
          }
          r = make_float2(A.x - B.y, A.y + B.x);
        }
        z[n1] = r;
      }
      reg_fft<N, 1>(z, twl, lds, ln, n2);
      if (ok) {
        const int y0 = 2 * p;
#pragma unroll
        for (int q = 0; q < V; ++q) {
          const int zz = n2 + L * q;
          if (zz < wz) {
            const unsigned lin0 = (unsigned)((xp * N + y0) * N + zz);
            const unsigned long long k0 = energy_key(z[q].x, lin0);
            best = k0 < best ? k0 : best;
            if (y0 + 1 < wy) {
              const unsigned long long k1 = energy_key(z[q].y, lin0 + (unsigned)N);
              best = k1 < best ? k1 : best;
            }
          }
        }
      }
    }
    __syncthreads();  // tile reads done before the next plane's writes
  }
  block_min_commit(best, keys + s);
}

// ---------------------------------------------------------------------------
// Column form of the fused pipeline (grid edge 64: 8 values per lane, two
// passes of 32 lines per workgroup).  The forward stage stops after the z transform; the mid stage runs,
// per kz plane, the y and x transforms of every channel through one LDS
// plane, the product with the protein spectrum, the channel sum, 1/G and the
// inverse x and y transforms; the last stage runs the z C2R and the argmin.
// Only the ligand box's (x, y) rows and the window's (x', y') rows cross HBM,
// about half the bytes of the plane form:
//   Z [slot][c][kz][n1 < R1][x < boxmax][n2], y = n2 + L*n1 < by, x < bx,
//     R1 = ceil(boxmax / L)
//   T2[slot][kz][m < V][x' < N][n2],          y' = n2 + L*m < wy, x' < wx
//   P2[c][kz][ky / LPW][q < V/2][n2][ky % LPW][2], kx = n2 + L*(2q + pair)
// Every index is a lane's register layout (line, n2, value n1), so each
// wavefront load or store covers 512 contiguous bytes (1 KB for P2).
// ---------------------------------------------------------------------------
// LDS plane of the column form's mid stage: (x, k) at x*N + (k ^ ((9x) & 15)).
// Lines along x with values along k, and lines along k with values along x,
// then both put every 16 consecutive lanes on 16 distinct 8-byte bank slots
// (checked by enumeration for N = 64; no unswizzled linear layout does).
// With L = 8 the swizzle of value j of a lane differs from value 0's only in
// bit 3, by j's parity, so a lane's V addresses are two bases plus
// compile-time offsets:
//   lane line x, value j at k = n2 + 8j:  base[j & 1] + 16 * (j >> 1)
//   lane line k, value j at x = n2 + 8j:  base[j & 1] + 8 * N * j
struct ColBases {
  int b0, b1;
};
template <int N>
__device__ __forceinline__ ColBases col_line_x(int x, int n2) {
  const int m = (9 * x) & 15;
  return {x * N + (n2 ^ m), x * N + ((n2 + 8) ^ m)};
}
template <int N>
__device__ __forceinline__ ColBases col_line_k(int k, int n2) {
  const int m = (9 * n2) & 15;
  return {n2 * N + (k ^ m), n2 * N + (k ^ m ^ 8)};
}
template <int N>
__device__ __forceinline__ int col_at_x(const ColBases& b, int j) {  // lane line x, value j
  return ((j & 1) ? b.b1 : b.b0) + 16 * (j >> 1);
}
template <int N>
__device__ __forceinline__ int col_at_k(const ColBases& b, int j) {  // lane line k, value j
  return ((j & 1) ? b.b1 : b.b0) + 8 * N * j;
}

// Position of spectrum value (kx, ky, kz) in a channel's P2 block.
template <int N>
__device__ __forceinline__ int kzb_index(int kx, int ky, int kz) {
  using RS = RegShape<N>;
  constexpr int L = RS::L, V = RS::V, LPW = RS::LPW, NB = N / LPW;
  const int m = kx / L, n2 = kx % L, kyb = ky / LPW, ln = ky % LPW;
  return (((kz * NB + kyb) * (V / 2) + m / 2) * L + n2) * (2 * LPW) + ln * 2 + (m & 1);
}

template <class P>
__global__ __launch_bounds__(kBlock) void k_to_kz_blocks(const float2* __restrict__ in, float2* __restrict__ out,
                                                         int nch) {
  constexpr int N = P::N, ZH = P::ZH;
  const size_t per = (size_t)N * N * ZH;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= per * nch) return;
  const int c = (int)(t / per);
  const int r = (int)(t - (size_t)c * per);  // [kx][ky][kz] (cuFFT R2C layout)
  const int kx = r / (N * ZH), ky = (r / ZH) % N, kz = r % ZH;
  out[(size_t)c * per + (size_t)kzb_index<N>(kx, ky, kz)] = in[t];
}

// Column form, forward: one workgroup per (pose, ligand channel); per
// occupied x-plane the register form's spread and z transforms
// (fwd_plane_rows), then the row pairs unpacked into half-spectrum rows
// y < by and stored kz-major into Z.
template <class P>
__global__ __launch_bounds__(kBlock) void k_fwd_z(const PoseSlot* __restrict__ slots,
                                                  const LigandDesc* __restrict__ ligs,
                                                  const int* __restrict__ chan_atoms, const float* __restrict__ wv,
                                                  const float* __restrict__ wq, const AtomGrid* __restrict__ ag,
                                                  int astride, const RotMeta* __restrict__ meta, int cmax,
                                                  int boxmax, int natom_lds, float2* __restrict__ Z,
                                                  const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  using RS_ = RegShape<N>;
  constexpr int L = RS_::L, V = RS_::V, RSTR = FwdRegShape<N>::RSTR;
  static_assert(V == 8 && L == 8, "column form: 8 lanes x 8 values per line");
  __shared__ float2 sbuf[4][RS_::SLICE];
  extern __shared__ double dyn[];  // [natom_lds] SpreadAtom, then the tile [npmax][RSTR]
  const int c = blockIdx.x, s = blockIdx.y;
  const RotMeta m = meta[s];
  if (!m.valid) return;
  const PoseSlot ps = slots[s];
  const LigandDesc& Ld = ligs[ps.lig];
  if (c >= Ld.nch) return;
  const int nc = Ld.ch_start[c + 1] - Ld.ch_start[c];
  SpreadAtom* sa = reinterpret_cast<SpreadAtom*>(dyn);
  float2* zt = reinterpret_cast<float2*>(sa + natom_lds);
  {
    const int* ca = chan_atoms + Ld.list_off + Ld.ch_start[c];
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
      const int a = ca[i];
      const AtomGrid g = ag[(size_t)s * astride + a];
      SpreadAtom t;
      t.ix = g.ix;
      t.iy = g.iy;
      t.iz = g.iz;
      t.w = channel_weight(Ld, c, a, wv, wq);
      t.fx = g.fx;
      t.fy = g.fy;
      t.fz = g.fz;
      sa[i] = t;
    }
  }
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int n2l = lane / RS_::LPW;
  float2 twr[V];
  reg_twiddles<N>(twg, n2l, twr);
  const int bxm = min(m.box[0], boxmax), by = m.box[1], bz = m.box[2];
  const int npair = (by + 1) >> 1;
  const bool zhalf = bz <= N / 2;
  const int r1 = (boxmax + L - 1) / L;
  float2* zb = Z + ((size_t)s * cmax + c) * ((size_t)ZH * r1 * boxmax * L);
  for (int x = 0; x < bxm; ++x) {
    fwd_plane_rows<N>(x, sa, nc, zt, npair, zhalf, twr, sbuf[wave], wave, lane);
    // thread t takes rows y = t % 32 (+ 32) of bins kz = t / 32 (+ 8 j):
    // power-of-two index math only
    for (int y = threadIdx.x & 31; y < by; y += 32) {
      const int p = y >> 1;
      float2* dst = zb + ((size_t)(y >> 3) * boxmax + x) * L + (y & 7);
      for (int kz = threadIdx.x >> 5; kz < ZH; kz += kBlock / 32) {
        const float2 zk = zt[p * RSTR + kz], zm = zt[p * RSTR + (kz == 0 ? 0 : N - kz)];
        const float2 v = (y & 1) ? make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x))
                                 : make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        dst[(size_t)kz * r1 * boxmax * L] = v;
      }
    }
  }
}

// Inverse part of the column-form mid stages, for one pose and kz plane:
// acc holds the product spectrum on the lane's ky lines (pass p line
// p*LPP + wave*LPW + ln, value m at kx = n2 + L*m).  1/G, inverse x
// transforms, back through the plane PROWS rows x' at a time, inverse y
// transforms of the window's x' lines, store of the window rows into tb
// ([m][x'][n2], y' = n2 + L*m).  Starts with a block barrier before its first
// plane write; all threads of the workgroup call it.  TR_IN_PLANE: the
// transforms' scratch tr lies inside the plane, so every transform waits
// until no wavefront reads the plane any more.
template <int N, int PROWS, bool TR_IN_PLANE = false>
__device__ __forceinline__ void col_inverse(float2 (&acc)[N / (4 * RegShape<N>::LPW)][RegShape<N>::V], float2* pl,
                                            float2* tr, const float2* twr, int lane, int wave, float2* tb, int wx,
                                            int wy, float invG) {
  using RS = RegShape<N>;
  constexpr int L = RS::L, V = RS::V, LPW = RS::LPW;
  constexpr int LPP = 4 * LPW;
  constexpr int NPASS = N / LPP;
  // inverse x transforms of the ky lines: value m = x' = n2 + L*m
  int ln = lane % LPW, n2 = lane / LPW;
  asm volatile("" : "+v"(ln), "+v"(n2));
  if constexpr (TR_IN_PLANE) __syncthreads();  // the plane's reads are done
#pragma unroll
  for (int p = 0; p < NPASS; ++p) {
#pragma unroll
    for (int i = 0; i < V; ++i) acc[p][i] = c_scale(acc[p][i], invG);
    reg_fft<N, 1>(acc[p], twr, tr, ln, n2);
  }
  // Back through the plane, N / ROUNDS rows x' at a time, then inverse y
  // transforms of the window's x' lines.
  constexpr int ROUNDS = N / PROWS;
  constexpr int QR = V / ROUNDS;  // row groups (values) per round
#pragma unroll
  for (int r = 0; r < ROUNDS; ++r) {
    __syncthreads();  // the plane's previous reads are done
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
      const ColBases wb = col_line_k<N>(p * LPP + wave * LPW + ln, n2);
#pragma unroll
      for (int q = 0; q < QR; ++q) pl[col_at_k<N>(wb, q)] = acc[p][r * QR + q];
    }
    __syncthreads();
    auto in_round = [&](int p) {
      const int x0 = p * LPP + wave * LPW;
      return x0 >= r * PROWS && x0 < (r + 1) * PROWS && x0 < wx;
    };
    auto store = [&](int p, const float2* v) {
      const int xp = p * LPP + wave * LPW + ln;
      if (xp < wx) {
#pragma unroll
        for (int q = 0; q < V; ++q)
          if (n2 + L * q < wy) tb[(q * N + xp) * L + n2] = v[q];
      }
    };
    if constexpr (TR_IN_PLANE) {
      // read every line first, then transform behind a barrier
      float2 v[NPASS][V];
#pragma unroll
      for (int p = 0; p < NPASS; ++p)
        if (in_round(p)) {
          const ColBases rb = col_line_x<N>(p * LPP + wave * LPW + ln - r * PROWS, n2);
#pragma unroll
          for (int n1 = 0; n1 < V; ++n1) v[p][n1] = pl[col_at_x<N>(rb, n1)];
        }
      __syncthreads();  // the plane's reads are done
#pragma unroll
      for (int p = 0; p < NPASS; ++p)
        if (in_round(p)) {
          reg_fft<N, 1>(v[p], twr, tr, ln, n2);
          store(p, v[p]);
        }
    } else {
#pragma unroll
      for (int p = 0; p < NPASS; ++p)
        if (in_round(p)) {
          float2 v[V];
          const ColBases rb = col_line_x<N>(p * LPP + wave * LPW + ln - r * PROWS, n2);
#pragma unroll
          for (int n1 = 0; n1 < V; ++n1) v[n1] = pl[col_at_x<N>(rb, n1)];
          reg_fft<N, 1>(v, twr, tr, ln, n2);
          store(p, v);
        }
    }
  }
}


// Column form, mid: one workgroup per (kz, group of poses).  For each pose
// and channel: y transforms of the box's x lines into the LDS plane [x][ky],
// x transforms of all ky lines out of it, product with conj and the protein
// spectrum, channel sum; then 1/G, inverse x transforms of the ky lines,
// back through the plane, inverse y transforms of the window's x' lines,
// store of the window rows.  Lane lines: 4 wavefronts x 64/L lines per
// pass, N / (4 * 64/L) passes.  When the box fits in N/2 along a transform's
// input axis the upper half is neither loaded nor transformed in the first
// DFT stage.  HALF: the kernel takes only the poses whose box fits in N/2
// along x, so the plane needs only N/2 rows (x < N/2 forward; the inverse
// passes it x' rows N/2 at a time) and four workgroups fit a CU; the full
// form takes the others (skip_small) or all.  Both round identically, so
// the form a pose goes through does not change its result.  Workgroup w
// runs on XCD w % 8 and takes kz = w % 8 + 8 j, so each XCD's L2 holds the
// protein spectra of every eighth kz plane only.
template <class P, bool HALF>
__global__ __launch_bounds__(kBlock) void k_mid_kz(const PoseSlot* __restrict__ slots,
                                                   const LigandDesc* __restrict__ ligs,
                                                   const RotMeta* __restrict__ meta, int nslots, int group,
                                                   const float2* __restrict__ Z, int cmax, int boxmax,
                                                   const float2* __restrict__ P2, float2* __restrict__ T2, float invG,
                                                   const float2* __restrict__ twg, int skip_small) {
  constexpr int N = P::N, ZH = P::ZH;
  using RS = RegShape<N>;
  constexpr int L = RS::L, V = RS::V, LPW = RS::LPW;
  static_assert(V == 8 && L == 8, "column form: 8 lanes x 8 values per line");
  constexpr int LPP = 4 * LPW;     // lines per pass
  static_assert(N % LPP == 0, "whole passes of lines");
  constexpr int NPASS = N / LPP;
  __shared__ float2 sbuf[4][RS::SLICE];
  constexpr int PROWS = HALF ? N / 2 : N;  // plane rows
  __shared__ float2 pl[PROWS * N];
  __shared__ float2 stw[L * V];
  int kz, sg;
  if (!xcd_unit_of(blockIdx.x, ZH / 8, ZH % 8, (nslots + group - 1) / group, kz, sg)) return;
  // twiddles from LDS: the registers go to the accumulators
  for (int t = threadIdx.x; t < L * V; t += blockDim.x) stw[t] = twg[(t / V) * (t % V)];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  float2* tr = sbuf[wave];
  const float2* twr = stw + (lane / LPW) * V;
  auto uni = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
  const int r1 = (boxmax + L - 1) / L;
  const size_t zslab = (size_t)ZH * r1 * boxmax * L;  // float2 per (pose, channel)
  const int s0 = sg * group, s1 = min(nslots, s0 + group);
  for (int s = s0; s < s1; ++s) {
    if (!uni(meta[s].valid)) continue;
    const int lig = uni(slots[s].lig);
    const int nch = uni(ligs[lig].nch);
    const int bx = uni(meta[s].box[0]), by = uni(meta[s].box[1]);
    if (HALF ? bx > N / 2 : (skip_small && bx <= N / 2)) continue;  // the other form's pose
    const int wx = uni(meta[s].win[0]), wy = uni(meta[s].win[1]);
    const bool xhalf = HALF || bx <= N / 2, yhalf = by <= N / 2;
    float2 acc[NPASS][V];
#pragma unroll
    for (int p = 0; p < NPASS; ++p)
#pragma unroll
      for (int i = 0; i < V; ++i) acc[p][i] = make_float2(0.f, 0.f);
    const int zrow = boxmax * L * (int)sizeof(float2);  // bytes per row group n1
    constexpr int NPY = HALF ? 1 : NPASS;                // passes of x lines (x < bx)
    // Z rows of channel c for the lines of pass p: the kz block of Z; lanes
    // outside the box read zero through the descriptor (offset past its
    // end), with no memory traffic.
    auto zfetch = [&](int c, int p, float2* dst) {
      int ln = lane % LPW, n2 = lane / LPW;
      asm volatile("" : "+v"(ln), "+v"(n2));
      const __amdgpu_buffer_rsrc_t zr =
          buffer_rsrc(Z + ((size_t)s * cmax + c) * zslab + (size_t)kz * r1 * boxmax * L, (unsigned)(r1 * zrow));
      const int x = p * LPP + wave * LPW + ln;
      const int voff = x < bx ? (x * L + n2) * (int)sizeof(float2) : kVoffNone;
#pragma unroll
      for (int n1 = 0; n1 < V; ++n1)
        dst[n1] = (n1 < V / 2 || !yhalf) ? buffer_load_f2(zr, n2 + L * n1 < by ? voff + n1 * zrow : kVoffNone)
                                         : make_float2(0.f, 0.f);
    };
    // The next channel's rows are in flight while this channel is transformed.
    float2 zp[NPY][V];
#pragma unroll
    for (int p = 0; p < NPY; ++p)
      if (p * LPP + wave * LPW < bx) zfetch(0, p, zp[p]);
    for (int c = 0; c < nch; ++c) {
      // Lane coordinates pass through an empty asm each iteration: the
      // address arithmetic below is then redone here instead of being
      // hoisted out of the loops into dozens of live registers.
      int ln = lane % LPW, n2 = lane / LPW;
      asm volatile("" : "+v"(ln), "+v"(n2));
      __syncthreads();  // the plane's previous reads are done
#pragma unroll
      for (int p = 0; p < NPY; ++p) {
        const int x0 = p * LPP + wave * LPW;
        if (x0 < bx) {
          const int x = x0 + ln;
          float2 v[V];
#pragma unroll
          for (int n1 = 0; n1 < V; ++n1) v[n1] = zp[p][n1];
          if (c + 1 < nch) zfetch(c + 1, p, zp[p]);
          if (yhalf)
            reg_fft<N, -1, true>(v, twr, tr, ln, n2);
          else
            reg_fft<N, -1, false>(v, twr, tr, ln, n2);
          const ColBases wb = col_line_x<N>(x, n2);
#pragma unroll
          for (int q = 0; q < V; ++q) pl[col_at_x<N>(wb, q)] = v[q];
        }
      }
      __syncthreads();
      const int pc = uni(ligs[lig].pmap[c]);
#pragma unroll
      for (int p = 0; p < NPASS; ++p) {
        const int ky0 = p * LPP + wave * LPW, ky = ky0 + ln;
        const float2* pb = P2 + ((size_t)pc * ZH + kz) * (size_t)(N * N) +
                           (size_t)(ky0 / LPW) * (V / 2) * L * 2 * LPW + n2 * 2 * LPW + ln * 2;
        float2 pr[V];
#pragma unroll
        for (int q = 0; q < V / 2; ++q) {
          const float4 f = *reinterpret_cast<const float4*>(pb + q * L * 2 * LPW);
          pr[2 * q] = make_float2(f.x, f.y);
          pr[2 * q + 1] = make_float2(f.z, f.w);
        }
        float2 v[V];
        const ColBases rb = col_line_k<N>(ky, n2);
        // rows x >= bx of the plane hold stale values: read, then select
        auto pread = [&](int n1) {
          const float2 t = pl[col_at_k<N>(rb, n1)];
          const bool in = n2 + L * n1 < bx;
          return make_float2(in ? t.x : 0.f, in ? t.y : 0.f);
        };
        if (xhalf) {
#pragma unroll
          for (int n1 = 0; n1 < V / 2; ++n1) v[n1] = pread(n1);
          reg_fft<N, -1, true>(v, twr, tr, ln, n2);
        } else {
#pragma unroll
          for (int n1 = 0; n1 < V; ++n1) v[n1] = pread(n1);
          reg_fft<N, -1, false>(v, twr, tr, ln, n2);
        }
#pragma unroll
        for (int i = 0; i < V; ++i) {
          const float2 l = v[i], q = pr[i];
          acc[p][i].x = fmaf(l.x, q.x, acc[p][i].x);
          acc[p][i].x = fmaf(l.y, q.y, acc[p][i].x);
          acc[p][i].y = fmaf(l.x, q.y, acc[p][i].y);
          acc[p][i].y = fmaf(-l.y, q.x, acc[p][i].y);
        }
        // keep the passes apart: overlapping them costs registers, not time
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    col_inverse<N, PROWS>(acc, pl, tr, twr, lane, wave, T2 + ((size_t)s * ZH + kz) * (size_t)(N * N), wx, wy, invG);
  }
}

// ---------------------------------------------------------------------------
// Atom form of the fused pipeline (grid edge 64).  Trilinear spreading is
// separable, so the DFT of a channel's ligand grid is, per kz plane, a sum of
// outer products over the channel's atoms:
//   L_c(kx, ky, kz) = sum_a w_a X_a(kx) Y_a(ky) Z_a(kz),
//   X_a(k) = (1 - fx) W^(k ix) + fx W^(k (ix + 1)),  W = exp(-2 pi i / N),
// (Y, Z likewise; w = sqrt(eps) or the charge, FillLigandGrid.cpp:45, :97).
// conj(L_c) on a kz plane is then a complex matrix product (64 kx) x (atoms)
// x (64 ky), run on the matrix cores: v_mfma_f32_32x32x2f32 with k = (atom,
// real/imaginary part), one instruction per atom for the real part of a
// 32 x 32 block and one for the imaginary part.  The ligand grids, their
// forward transforms and the forward stage's HBM traffic disappear; the
// product with the protein spectrum, the channel sum and the inverse x and y
// transforms follow as in the column form, and k_inv_z finishes.
//
// MFMA 32x32x2 f32 lane layout (checked on the device): lane l supplies
// A[l % 32][l / 32] and B[l / 32][l % 32]; result register r of lane l is
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
// P3[c][kz][block b = (N/32) (kx / 32) + ky / 32][r / 2][l][2]: the protein
// spectra in that result layout (one 16-byte load gives a lane registers
// 2q and 2q + 1; a wavefront load reads 1 KB).
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ int mfma_index(int kx, int ky, int kz) {
  constexpr int NB = N / 32;  // 32 x 32 blocks per row of the plane
  const int row = kx & 31, col = ky & 31;
  const int l = col + 32 * ((row >> 2) & 1), r = 4 * (row >> 3) + (row & 3);
  const int b = NB * (kx >> 5) + (ky >> 5);
  return (((kz * NB * NB + b) * 8 + (r >> 1)) * 64 + l) * 2 + (r & 1);
}

template <class P>
__global__ __launch_bounds__(kBlock) void k_to_mfma_blocks(const float2* __restrict__ in, float2* __restrict__ out,
                                                           int nch) {
  constexpr int N = P::N, ZH = P::ZH;
  const size_t per = (size_t)N * N * ZH;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= per * nch) return;
  const int c = (int)(t / per);
  const int r = (int)(t - (size_t)c * per);  // [kx][ky][kz] (cuFFT R2C layout)
  const int kx = r / (N * ZH), ky = (r / ZH) % N, kz = r % ZH;
  out[(size_t)c * per + (size_t)mfma_index<N>(kx, ky, kz)] = in[t];
}

typedef float mfma_acc16 __attribute__((ext_vector_type(16)));
#ifndef CDK_ATOM_GROUP
#define CDK_ATOM_GROUP 4
#endif

typedef _Float16 mfma_h8 __attribute__((ext_vector_type(8)));
typedef unsigned mfma_u4 __attribute__((ext_vector_type(4)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));

// Full-rate 24-bit multiply (the compiler narrows __umul24 to a quarter-rate
// 32-bit multiply when only the product's low bits are used).
__device__ __forceinline__ unsigned mul_u24(unsigned a, unsigned b) {
  unsigned r;
  asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// Byte offset of W^(k i) in a table of N float2 (i = 8 * index at N = 64).
template <int N>
__device__ __forceinline__ unsigned atom_tw_bytes(unsigned k, unsigned i);

// Stages padded slots [t_lo, t_hi) of one pose's channel lists for the atom
// form (LigandDesc::pad_atoms: each channel padded to a multiple of 8 entries
// by -1) into LDS slots 0..: a listed atom as ef = (wZ, fx wZ), wZ = the
// weight times Z_a(kz) = (1 - fz) W^(kz iz) + fz W^(kz (iz + 1)), and ei =
// (ix, iy, fy bits, 0), ix and iy times 8 (table byte offsets) at grid edges
// 64 and 128; a pad as zeros (zero weight: every product it enters is exactly 0).
// ag: the pose's placements.  SC (the FP16 split products): the weight is
// scaled by Ld.ascale and fy by kAtomYScale (powers of two, exact).
template <int N, bool SC>
__device__ void stage_atoms(float4* ef, int4* ei, const LigandDesc& Ld, int t_lo, int t_hi,
                            const int* __restrict__ pad_atoms, const float* __restrict__ wv,
                            const float* __restrict__ wq, const AtomGrid* __restrict__ ag, int kz,
                            const float2* stw) {
  const int pad_off = Ld.pad_off, w_off = Ld.w_off, t_elec = Ld.pad_start[Ld.nch - 1];
  __syncthreads();  // the previous staged atoms (and anything sharing their memory) are done with
  for (int k = threadIdx.x; k < t_hi - t_lo; k += blockDim.x) {
    const int t = t_lo + k;
    const int a = pad_atoms[pad_off + t];
    float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
    int4 q = make_int4(0, 0, 0, 0);
    if (a >= 0) {
      const AtomGrid g = ag[a];
      // elec channel: the charge; vdW channels: sqrt(eps) (FillLigandGrid.cpp:45, :97)
      const float wa = (t >= t_elec ? wq[w_off + a] : wv[w_off + a]) * (SC ? Ld.ascale : 1.f);
      const float2 z0 = stw[(kz * g.iz) % N], z1 = stw[(kz * (g.iz + 1)) % N];
      const float2 zf = c_fma_s(z1, (float)g.fz, c_scale(z0, (float)(1.0 - g.fz)));
      const float2 wz = make_float2(wa * zf.x, wa * zf.y);
      const float fx = (float)g.fx;
      constexpr int S = (N & (N - 1)) == 0 ? 8 : 1;  // byte offsets for a power-of-two N
      f = make_float4(wz.x, wz.y, fx * wz.x, fx * wz.y);
      q = make_int4(S * g.ix, S * g.iy, __float_as_int((float)g.fy * (SC ? kAtomYScale : 1.f)), 0);
    }
    ef[k] = f;
    ei[k] = q;
  }
  __syncthreads();
}

// Per pose: stages the whole padded list when it fits `ent` slots, else one
// channel at a time (a channel pads to at most kMaxAtoms entries), and runs
// body(c, k0, n) per channel with its n entries at staged slots [k0, k0 + n)
// and zero entries after them to the next multiple of 8.
template <int N, bool SC, class Body>
__device__ __forceinline__ void for_staged_channels(float4* ef, int4* ei, int ent, const LigandDesc& Ld,
                                                    const int* __restrict__ pad_atoms,
                                                    const float* __restrict__ wv, const float* __restrict__ wq,
                                                    const AtomGrid* __restrict__ ag, int kz, const float2* stw,
                                                    Body&& body) {
  auto uni = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
  const int nch = uni(Ld.nch);
  const int ptot = uni(Ld.pad_start[nch]);
  const bool whole = ptot <= ent;
  if (whole) stage_atoms<N, SC>(ef, ei, Ld, 0, ptot, pad_atoms, wv, wq, ag, kz, stw);
  for (int c = 0; c < nch; ++c) {
    const int n = uni(Ld.ch_start[c + 1]) - uni(Ld.ch_start[c]);
    const int p0 = uni(Ld.pad_start[c]);
    if (!whole) stage_atoms<N, SC>(ef, ei, Ld, p0, uni(Ld.pad_start[c + 1]), pad_atoms, wv, wq, ag, kz, stw);
    body(c, whole ? p0 : 0, n);
  }
}

// x = hi + lo to within 2^-22 |x| (normal range): hi = f16(x), lo = f16(x -
// hi), both rounded to nearest, two values per register.  The residual x - hi
// (exact in FP32) comes from one mixed-precision fma per value.
__device__ __forceinline__ void split_f16(float a, float b, unsigned& hi, unsigned& lo) {
  hi = __builtin_bit_cast(unsigned, __builtin_convertvector((float2v){a, b}, half2v));
  float ra, rb;
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(ra) : "v"(hi), "v"(a));                // a - hi.lo
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(rb) : "v"(hi), "v"(b));  // b - hi.hi
  lo = __builtin_bit_cast(unsigned, __builtin_convertvector((float2v){ra, rb}, half2v));
}

// x = hi + mid + lo exactly (FP32 x of magnitude in [2^-1, 2^15); below
// that lo rounds to the FP16 subnormal spacing 2^-24): hi = f16(x), mid =
// f16(x - hi), lo = f16(x - hi - mid), both residuals exact in FP32 (x - hi
// has at most 13 significant bits, x - hi - mid at most 2), each from one
// mixed-precision fma per value; two values per register.
__device__ __forceinline__ void split3_f16(float a, float b, unsigned& hi, unsigned& mid, unsigned& lo) {
  hi = __builtin_bit_cast(unsigned, __builtin_convertvector((float2v){a, b}, half2v));
  float ra, rb, sa, sb;
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(ra) : "v"(hi), "v"(a));                // a - hi.lo
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(rb) : "v"(hi), "v"(b));  // b - hi.hi
  mid = __builtin_bit_cast(unsigned, __builtin_convertvector((float2v){ra, rb}, half2v));
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(sa) : "v"(mid), "v"(ra));
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(sb) : "v"(mid), "v"(rb));
  lo = __builtin_bit_cast(unsigned, __builtin_convertvector((float2v){sa, sb}, half2v));
}

// (x, y) -> (-y, x) on a packed FP16 pair (an exact multiply by 1).
__device__ __forceinline__ unsigned swap_neg_f16(unsigned h) {
  unsigned r;
  asm("v_pk_mul_f16 %0, %1, 1.0 op_sel:[1,0] op_sel_hi:[0,0] neg_lo:[1,0]" : "=v"(r) : "v"(h));
  return r;
}

// A operands shared by the SW wavefronts of a workgroup that own blocks of the
// same kx rows (SW = 1: each wavefront builds its own): per 8-entry step,
// wavefront w builds and splits the A values of entries j = w, w + SW, ...
// (j < 4 per lane half) into buf[par][j][lane] = (hi, lo), one barrier, and
// every wavefront reads all four; two buffers alternate, so one barrier per
// step suffices (a buffer is rewritten only after every wavefront passed the
// barrier that follows its last reads).
struct AShare {
  uint4* buf;  // [2][4][64]: (hi, lo) or (hi, mid, lo)
  int wave;    // this wavefront's index among the SW
  int par;     // buffer of the next step
};

// One channel's conj(L_c) on this lane's 32 x 32 block (cr, ci: real and
// imaginary parts, written here): entries [k0, kend)
// of the staged lists ef and ei (zero entries follow to the next multiple of
// 8 after kend), with
// A[kx][(a, part)] = the part of conj(w Z X_a(kx)) and B[(a, part)][ky] =
// the parts of conj(Y_a(ky)) that give the real (cr) and imaginary (ci)
// sums, X_a(k) = W^(k ix) (1 - fx D(k)), Y_a(k) = W^(k iy) (1 - fy D(k)),
// D = 1 - W^k (dx, dy: this lane's D at kxA, kyB).
//
// PR = kAtomF16 / kAtomF16x3: the products on the FP16 matrix cores (32 x 32
// x 16: eight atoms per instruction, lane half h builds both parts of atoms
// 4h..4h+3 of each eight).  The staged weights and fy carry the power-of-two
// scales (stage_atoms), so A is scaled by ascale and B by kAtomYScale.
// kAtomF16: every operand split x = hi + lo and A B = Ah Bh + Ah Bl + Al Bh
// accumulated in FP32 (the dropped Al Bl and the splits leave about 2^-21 per
// product).  kAtomF16x3: x = hi + mid + lo exactly and the six partial
// products down to 2^-22 (Ah Bh, Ah Bm, Am Bh, Ah Bl, Al Bh, Am Bm; the
// dropped ones are below 2^-32), each FP16 product exact in FP32.  Slots past
// kend are the staged zero entries.  kAtomF32: FP32 matrix cores (32 x 32 x
// 2: one atom per instruction, lane half = part), AU atoms' operands in
// flight at once, no scales.
template <int N, int PR, int SW = 1>
__device__ __forceinline__ void atom_channel(const float4* ef, const int4* ei, const char* stwb, int k0, int kend,
                                             int kxA, int kyB, float2 dx, float2 dy, int part, mfma_acc16& cr,
                                             mfma_acc16& ci, AShare* share = nullptr) {
  constexpr bool F16 = PR != kAtomF32;
  constexpr bool X3 = PR == kAtomF16x3;
  auto lookups = [&](int k, float4& f, float2& tx, float2& ty, float& fy) {
    f = ef[k];
    const int4 ii = ei[k];
    tx = *reinterpret_cast<const float2*>(stwb + atom_tw_bytes<N>(kxA, ii.x));
    ty = *reinterpret_cast<const float2*>(stwb + atom_tw_bytes<N>(kyB, ii.y));
    fy = __int_as_float(ii.z);
  };
  if constexpr (F16) {
    // (a0, a1) = conj(W^(kx ix) g), g = wZ - fx wZ D, k slots (real, imaginary)
    auto a_vals = [&](int kk, float& a0, float& a1) {
      const float4 f = ef[kk];
      const float2 tx = *reinterpret_cast<const float2*>(stwb + atom_tw_bytes<N>(kxA, ei[kk].x));
      const float gx = fmaf(-f.z, dx.x, fmaf(f.w, dx.y, f.x));
      const float gy = fmaf(-f.z, dx.y, fmaf(-f.w, dx.x, f.y));
      a0 = fmaf(tx.x, gx, -(tx.y * gy));
      a1 = fmaf(-tx.x, gy, -(tx.y * gx));
    };
    auto step = [&](int k, auto first) {
      // parts: [0] hi, [1] lo (kAtomF16) or [1] mid, [2] lo (kAtomF16x3)
      constexpr int NP = X3 ? 3 : 2;
      unsigned a[NP][4], b[NP][4];
      if constexpr (SW > 1) {
        uint4* sb = share->buf + share->par * (4 * 64);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (j % SW == share->wave) {
            float a0, a1;
            a_vals(k + 4 * part + j, a0, a1);
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if constexpr (X3)
              split3_f16(a0, a1, v.x, v.y, v.z);
            else
              split_f16(a0, a1, v.x, v.y);
            sb[j * 64 + (threadIdx.x & 63)] = v;
          }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int kk = k + 4 * part + j;
        float2 ty;
        float fy, a0 = 0.f, a1 = 0.f;
        if constexpr (SW == 1) {
          // this lane's A and B for entry kk, lookups together
          float4 f;
          float2 tx;
          lookups(kk, f, tx, ty, fy);
          const float gx = fmaf(-f.z, dx.x, fmaf(f.w, dx.y, f.x));
          const float gy = fmaf(-f.z, dx.y, fmaf(-f.w, dx.x, f.y));
          a0 = fmaf(tx.x, gx, -(tx.y * gy));
          a1 = fmaf(-tx.x, gy, -(tx.y * gx));
        } else {
          const int4 ii = ei[kk];
          ty = *reinterpret_cast<const float2*>(stwb + atom_tw_bytes<N>(kyB, ii.y));
          fy = __int_as_float(ii.z);
        }
        // kAtomYScale (1 - fy D), fy staged times kAtomYScale
        const float hx = fmaf(-fy, dy.x, kAtomYScale), hy = -fy * dy.y;
        const float yx = fmaf(ty.x, hx, -(ty.y * hy)), yy = fmaf(ty.x, hy, ty.y * hx);
        if constexpr (X3) {
          if constexpr (SW == 1) split3_f16(a0, a1, a[0][j], a[1][j], a[2][j]);
          split3_f16(yx, yy, b[0][j], b[1][j], b[2][j]);  // real sum: a0 yx + a1 yy
        } else {
          if constexpr (SW == 1) split_f16(a0, a1, a[0][j], a[1][j]);
          split_f16(yx, yy, b[0][j], b[1][j]);
        }
      }
      if constexpr (SW > 1) {
        __syncthreads();
        const uint4* sb = share->buf + share->par * (4 * 64);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint4 v = sb[j * 64 + (threadIdx.x & 63)];
          a[0][j] = v.x;
          a[1][j] = v.y;
          if constexpr (X3) a[NP - 1][j] = v.z;
        }
        share->par ^= 1;
      }
      auto op = [](const unsigned (&r)[4]) { return __builtin_bit_cast(mfma_h8, (mfma_u4){r[0], r[1], r[2], r[3]}); };
      // (i, j) part pairs of the kept partial products, largest first
      auto run = [&](mfma_acc16& acc, const unsigned (&bb)[NP][4]) {
        if constexpr (decltype(first)::value) {
          const mfma_acc16 zero = {};
          acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(op(a[0]), op(bb[0]), zero, 0, 0, 0);
        } else {
          acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(op(a[0]), op(bb[0]), acc, 0, 0, 0);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(op(a[0]), op(bb[1]), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(op(a[1]), op(bb[0]), acc, 0, 0, 0);
        if constexpr (X3) {
          acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(op(a[0]), op(bb[2]), acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(op(a[2]), op(bb[0]), acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(op(a[1]), op(bb[1]), acc, 0, 0, 0);
        }
      };
      run(cr, b);
      // imaginary sum: -a0 yy + a1 yx, B parts swapped to (-yy, yx)
#pragma unroll
      for (int i = 0; i < NP; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) b[i][j] = swap_neg_f16(b[i][j]);
      run(ci, b);
    };
    // At 96 and 128 the first step starts from a zero accumulator operand (no
    // zeroing moves); the fused 64-point kernel, at its register limit, runs
    // faster with one copy of the step.
    if (N == 64 || k0 >= kend) {
      // 64-bit moves: the zeroing takes half the instructions the compiler's
      // own 32-bit copies of one zero register take here
      typedef double zd8 __attribute__((ext_vector_type(8)));
      zd8 zr, zi;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        double a, b;
        asm volatile("v_mov_b64 %0, 0" : "=v"(a));
        asm volatile("v_mov_b64 %0, 0" : "=v"(b));
        zr[r] = a;
        zi[r] = b;
      }
      cr = __builtin_bit_cast(mfma_acc16, zr);
      ci = __builtin_bit_cast(mfma_acc16, zi);
    }
    if constexpr (N == 64) {
      for (int k = k0; k < kend; k += 8) step(k, std::false_type{});
    } else if (k0 < kend) {
      step(k0, std::true_type{});
      for (int k = k0 + 8; k < kend; k += 8) step(k, std::false_type{});
    }
  } else {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      cr[r] = 0.f;
      ci[r] = 0.f;
    }
    constexpr int AU = CDK_ATOM_GROUP;
    // A = conj(w Z X) = conj(W^(kx ix) (wZ - fx wZ D)) (one component per
    // lane: k slot `part`), B = conj(Y) = conj(W^(ky iy) (1 - fy D)).
    auto atom = [&](int k, float& av, float& br, float& bi) {
      float4 f;
      float2 tx, ty;
      float fy;
      lookups(k, f, tx, ty, fy);
      const float gx = fmaf(-f.z, dx.x, fmaf(f.w, dx.y, f.x));
      const float gy = fmaf(-f.z, dx.y, fmaf(-f.w, dx.x, f.y));
      const float p1 = part ? -gy : gx, p2 = part ? -gx : -gy;  // (A.x) or (-A.y) of tx g
      av = fmaf(tx.x, p1, tx.y * p2);
      const float hx = fmaf(-fy, dy.x, 1.f), hy = -fy * dy.y;
      const float yx = fmaf(ty.x, hx, -(ty.y * hy)), yy = fmaf(ty.x, hy, ty.y * hx);
      // conj(Y) = (yx, -yy): real part Br Ar - Bi Ai, imaginary Bi Ar + Br Ai
      br = part ? yy : yx;
      bi = part ? yx : -yy;
    };
    int k = k0;
    for (; k + AU <= kend; k += AU) {
      float av[AU], br[AU], bi[AU];
#pragma unroll
      for (int u = 0; u < AU; ++u) atom(k + u, av[u], br[u], bi[u]);
#pragma unroll
      for (int u = 0; u < AU; ++u) {
        cr = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u], br[u], cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u], bi[u], ci, 0, 0, 0);
      }
    }
    for (; k < kend; ++k) {
      float av, br, bi;
      atom(k, av, br, bi);
      cr = __builtin_amdgcn_mfma_f32_32x32x2f32(av, br, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bi, ci, 0, 0, 0);
    }
  }
}

template <>
__device__ __forceinline__ unsigned atom_tw_bytes<64>(unsigned k, unsigned i) {
  return mul_u24(k, i) & (8u * 64u - 8u);
}

template <>
__device__ __forceinline__ unsigned atom_tw_bytes<128>(unsigned k, unsigned i) {
  return mul_u24(k, i) & (8u * 128u - 8u);
}

// Atom form, mid: one workgroup per (kz, group of poses); wavefront w owns
// the 32 x 32 (kx, ky) block (w / 2, w % 2).  Per pose the ligand's channel
// atoms (channel order) are staged in LDS, all of them when they fit, else
// one channel at a time: wZ = the weight times Z_a(kz), fx wZ, ix, iy, fy.  Per
// channel every wavefront walks the channel's atoms in order,
// builds its A and B operands from the twiddle table (the next atom's
// lookups in flight while the current one is multiplied) and accumulates
// conj(L_c) on the matrix cores; at each channel end it multiplies by the
// protein spectrum into the pose's sum.
// The sum goes through the LDS plane into the ky-line register layout and
// col_inverse finishes.  Per pose the arithmetic order is fixed (atoms in
// channel order, one accumulation chain per channel), so results do not
// depend on the group size or on what else shares the launch.
template <class P, int PR>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_mid_atoms(const PoseSlot* __restrict__ slots,
                                                      const LigandDesc* __restrict__ ligs,
                                                      const int* __restrict__ pad_atoms,
                                                      const float* __restrict__ wv, const float* __restrict__ wq,
                                                      const AtomGrid* __restrict__ ag, int astride,
                                                      const RotMeta* __restrict__ meta, int nslots, int group,
                                                      const float2* __restrict__ P3, float2* __restrict__ T2,
                                                      float invG, const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  using RS = RegShape<N>;
  constexpr int L = RS::L, V = RS::V, LPW = RS::LPW;
  static_assert(N == 64 && V == 8 && L == 8, "atom form: 64 points, 8 lanes x 8 values per line");
  constexpr int LPP = 4 * LPW;
  constexpr int NPASS = N / LPP;
  constexpr int ENT = 512;  // channel atoms staged at a time
  // The staged atoms, the plane's contents and the inverse transforms'
  // scratch are never live together: all three share the plane's memory.
  constexpr int SCR = 4 * RS::SLICE * (int)sizeof(float2);
  constexpr int STG = ENT * (int)(2 * sizeof(float4));
  static_assert(SCR <= N * N * (int)sizeof(float2) && STG <= N * N * (int)sizeof(float2), "fits the plane");
  __shared__ __align__(16) float2 pl[N * N];
  unsigned char* scratch = reinterpret_cast<unsigned char*>(pl);
  __shared__ float2 stw[N];       // W^m
  __shared__ float2 stwr[L * V];  // register FFT twiddles
  float4* ef = reinterpret_cast<float4*>(scratch);                      // (w Z(kz), fx w Z(kz))
  int4* ei = reinterpret_cast<int4*>(scratch + ENT * sizeof(float4));  // (8 ix, 8 iy, fy bits, 0)
  int kz, sg;
  if (!xcd_plane_of(blockIdx.x, ZH / 8, ZH % 8, (nslots + group - 1) / group, kz, sg)) return;
  const int tid = threadIdx.x;
  for (int t = tid; t < N; t += blockDim.x) stw[t] = twg[t];
  for (int t = tid; t < L * V; t += blockDim.x) stwr[t] = twg[(t / V) * (t % V)];
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int part = lane >> 5;                      // k slot: real (0) or imaginary (1) part
  const int kxA = 32 * (wave >> 1) + (lane & 31);  // this lane's A row
  const int kyB = 32 * (wave & 1) + (lane & 31);   // this lane's B column
  float2* tr = reinterpret_cast<float2*>(scratch) + wave * RS::SLICE;
  const float2* twr = stwr + (lane / LPW) * V;
  auto uni = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
  __syncthreads();
  // X_a(k) = W^(k ix) (1 - fx (1 - W^k)): per lane, D = 1 - W^k
  const float2 dx = make_float2(1.f - stw[kxA].x, -stw[kxA].y), dy = make_float2(1.f - stw[kyB].x, -stw[kyB].y);
  const char* stwb = reinterpret_cast<const char*>(stw);
  const int s0 = sg * group, s1 = min(nslots, s0 + group);
  for (int s = s0; s < s1; ++s) {
    if (!uni(meta[s].valid)) continue;
    const int lig = uni(slots[s].lig);
    const LigandDesc& Ld = ligs[lig];
    const int wx = uni(meta[s].win[0]), wy = uni(meta[s].win[1]);
    float2 E[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) E[r] = make_float2(0.f, 0.f);
    for_staged_channels<N, PR != kAtomF32>(ef, ei, ENT, Ld, pad_atoms, wv, wq, ag + (size_t)s * astride, kz, stw,
                                           [&](int c, int k0, int n) {
      mfma_acc16 cr, ci;
      atom_channel<N, PR>(ef, ei, stwb, k0, k0 + n, kxA, kyB, dx, dy, part, cr, ci);
      // sum += conj(L_c) P_c
      const int pc = uni(Ld.pmap[c]);
      const float2* pb = P3 + (((size_t)pc * ZH + kz) * 4 + wave) * (8 * 128) + lane * 2;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float4 pq = *reinterpret_cast<const float4*>(pb + q * 128);  // registers 2q, 2q + 1
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int r = 2 * q + h;
          const float px = h ? pq.z : pq.x, py = h ? pq.w : pq.y;
          E[r].x = fmaf(cr[r], px, E[r].x);
          E[r].x = fmaf(-ci[r], py, E[r].x);
          E[r].y = fmaf(cr[r], py, E[r].y);
          E[r].y = fmaf(ci[r], px, E[r].y);
        }
      }
    });
    // the sum through the plane into the ky-line register layout
    __syncthreads();  // the staged atoms' reads are done
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int kx = 32 * (wave >> 1) + 8 * (r >> 2) + 4 * part + (r & 3);
      pl[kx * N + (kyB ^ ((9 * kx) & 15))] = E[r];
    }
    __syncthreads();
    float2 acc[NPASS][V];
    {
      int ln = lane % LPW, n2 = lane / LPW;
      asm volatile("" : "+v"(ln), "+v"(n2));
#pragma unroll
      for (int p = 0; p < NPASS; ++p) {
        const ColBases rb = col_line_k<N>(p * LPP + wave * LPW + ln, n2);
#pragma unroll
        for (int j = 0; j < V; ++j) acc[p][j] = pl[col_at_k<N>(rb, j)];
      }
    }
    // the split products' scales out with 1/G (powers of two: exact)
    const float sc = PR == kAtomF32 ? invG : invG * __int_as_float(uni(__float_as_int(Ld.aunscale))) * (1.f / kAtomYScale);
    col_inverse<N, N, true>(acc, pl, tr, twr, lane, wave, T2 + ((size_t)s * ZH + kz) * (size_t)(N * N), wx, wy, sc);
  }
}

// Atom form at grid edges 96 and 128, in two kernels (a 96- or 128-point
// line holds 24 or 16 values per lane, so a whole plane does not fit one
// workgroup's registers as at 64).
//
// k_atoms_fwd: one workgroup per (kz, row of 32 x 32 kx blocks, group of
// poses), N / 32 wavefronts, one block each: the same matrix-core ligand
// spectra and product as k_mid_atoms (at 96 the table index reduced mod 96
// exactly by a multiply and shift: (t * 43691) >> 22 = t / 96 for t <= 95 *
// 95).  The row's wavefronts own the same kx rows and share the A operands.
// Per pose the row's product spectrum (32 kx rows x N ky, times 1/G) goes
// into LDS tiles of 16 kx rows, the inverse y transforms run as LDS
// Stockham passes (few registers: the products keep the register file),
// and the window's y rows go to R[slot][y][kz][kx].
//
// k_atoms_xz: one workgroup per (pose, window row y), N / 2 / (lines per
// wavefront) wavefronts (3 at 96, 8 at 128): the inverse x transforms of the
// row's N/2 + 1 kz lines (the kz = 0 and N/2 lines as one complex line
// A + iB: both planes are Hermitian in (kx, ky), so after the x and y
// transforms both are real, as the z C2R takes them), the window's x'
// values through LDS into z lines, the z C2R of two x' lines at a time as
// one complex line (two-for-one, imaginary parts of the kz = 0 and N/2 bins
// dropped as a C2R transform does), the window minimum, one atomicMin per
// workgroup.  The product spectra make one HBM round trip, of the window's
// y rows only.
template <>
__device__ __forceinline__ unsigned atom_tw_bytes<96>(unsigned k, unsigned i) {
  const unsigned t = mul_u24(k, i);
  const unsigned q = mul_u24(t, 43691u) >> 22;
  return (t - 96u * q) * 8u;
}

template <int N>
struct AtomRows {
  using RS = RegShape<N>;
  static constexpr int ENT = 512;           // staged channel-list entries
  // Row stride of the LDS tiles (float2), odd: consecutive lines (the
  // Stockham passes' consecutive threads, and the x-z kernel's lines ln of
  // one 8-byte access) fall on distinct bank pairs.
  static constexpr int TS = N + 1;
  static constexpr int STAGE = ENT * 4 + 2 * 4 * 64 * 2;  // ef, ei and the shared operands (float2)
  static constexpr int TILES = 2 * 16 * TS;               // two tiles of 16 lines (Stockham ping-pong)
  static constexpr int U = STAGE > TILES ? STAGE : TILES;
  static_assert(U % 2 == 0, "float4 storage");
  // k_atoms_xz
  static constexpr int XW = N / 2 / RS::LPW;     // wavefronts
  static constexpr int XRW = N == 96 ? 2 : XW;   // of them transforming at a time
  static constexpr int XT = (N / 2 + 1) * TS;    // [kz][x'] tile
  static constexpr int XU = XT > XRW * RS::SLICE ? XT : XRW * RS::SLICE;
  // wavefronts per SIMD the tile's LDS allows (160 KB per CU): the compiler
  // may spend registers up to that occupancy on loads in flight
  static constexpr int XOCC = (160 * 1024 / (XU * 8 + 2048)) * XW / 4;
};

template <class P, int PR>
__global__ __launch_bounds__(256) void k_atoms_fwd(const PoseSlot* __restrict__ slots,
                                                     const LigandDesc* __restrict__ ligs,
                                                     const int* __restrict__ pad_atoms,
                                                     const float* __restrict__ wv, const float* __restrict__ wq,
                                                     const AtomGrid* __restrict__ ag, int astride,
                                                     const RotMeta* __restrict__ meta, int nslots, int group,
                                                     const float2* __restrict__ P3, float2* __restrict__ R,
                                                     float invG, const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  static_assert(N == 96 || N == 128, "the two-kernel atom form: grid edges 96 and 128");
  using AR = AtomRows<N>;
  constexpr int NB = N / 32, NW = NB * NB;  // 32 x 32 blocks per plane
  constexpr int ENT = AR::ENT;
  __shared__ float2 stw[N];
  // The staged atoms and shared operands and the transform tiles are never
  // live together.
  __shared__ float4 ubuf4[AR::U / 2];
  float2* ubuf = reinterpret_cast<float2*>(ubuf4);
  float4* ef = ubuf4;                               // (w Z(kz), fx w Z(kz))
  int4* ei = reinterpret_cast<int4*>(ubuf4 + ENT);  // (ix, iy, fy bits, 0; times 8 at 128)
  uint4* abuf = reinterpret_cast<uint4*>(ubuf4 + 2 * ENT);  // [2][4][64] shared A operands
  int kz, unit;
  if (!xcd_plane_of(blockIdx.x, ZH / 8, ZH % 8, NB * ((nslots + group - 1) / group), kz, unit)) return;
  const int brow = unit % NB, sg = unit / NB;
  const int tid = threadIdx.x;
  for (int t = tid; t < N; t += blockDim.x) stw[t] = twg[t];
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int part = lane >> 5;
  const int blk = NB * brow + wave;
  const int kxA = 32 * brow + (lane & 31);
  const int kyB = 32 * wave + (lane & 31);
  const char* stwb = reinterpret_cast<const char*>(stw);
  auto uni = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
  __syncthreads();
  const float2 dx = make_float2(1.f - stw[kxA].x, -stw[kxA].y), dy = make_float2(1.f - stw[kyB].x, -stw[kyB].y);
  AShare share{abuf, wave, 0};
  const int s0 = sg * group, s1 = min(nslots, s0 + group);
  for (int s = s0; s < s1; ++s) {
    if (!uni(meta[s].valid)) continue;
    const int lig = uni(slots[s].lig);
    const LigandDesc& Ld = ligs[lig];
    float2 E[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) E[r] = make_float2(0.f, 0.f);
    for_staged_channels<N, PR != kAtomF32>(ef, ei, ENT, Ld, pad_atoms, wv, wq, ag + (size_t)s * astride, kz, stw,
                                           [&](int c, int k0, int n) {
      mfma_acc16 cr, ci;
      atom_channel<N, PR, PR != kAtomF32 ? NB : 1>(ef, ei, stwb, k0, k0 + n, kxA, kyB, dx, dy, part, cr, ci, &share);
      // sum += conj(L_c) P_c
      const int pc = uni(Ld.pmap[c]);
      const float2* pb = P3 + (((size_t)pc * ZH + kz) * NW + blk) * (8 * 128) + lane * 2;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float4 pq = *reinterpret_cast<const float4*>(pb + q * 128);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int r = 2 * q + h;
          const float px = h ? pq.z : pq.x, py = h ? pq.w : pq.y;
          E[r].x = fmaf(cr[r], px, E[r].x);
          E[r].x = fmaf(-ci[r], py, E[r].x);
          E[r].y = fmaf(cr[r], py, E[r].y);
          E[r].y = fmaf(ci[r], px, E[r].y);
        }
      }
    });
    // 1/G and the split products' scales out (powers of two: exact) in one factor
    float sc = invG;
    if constexpr (PR != kAtomF32)
      sc = invG * (__int_as_float(uni(__float_as_int(Ld.aunscale))) * (1.f / kAtomYScale));
    const int wy = uni(meta[s].win[1]);
    float2* rb = R + (size_t)s * ((size_t)ZH * N * N) + (size_t)kz * N + 32 * brow;
    // the row's 32 kx rows in two halves of 16 ky lines: into the tile,
    // inverse y transforms through LDS, window rows out (consecutive threads
    // on consecutive kx)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      __syncthreads();  // the staged atoms and shared operands, or the previous half, are done with
      {
        float2* tr = ubuf + kyB;
#pragma unroll
        for (int r = 8 * h; r < 8 * h + 8; ++r)
          tr[(8 * (r >> 2) + 4 * part + (r & 3) - 16 * h) * AR::TS] = c_scale(E[r], sc);
      }
      __syncthreads();
      const float2* res = lds_fft<P, 1, true, 16>(ubuf, ubuf + 16 * AR::TS, 16, 1, AR::TS, stw);
      for (int t = tid; t < 16 * wy; t += blockDim.x) {
        const int line = t & 15, y = t >> 4;
        rb[(size_t)y * (ZH * N) + 16 * h + line] = res[line * AR::TS + y];
      }
    }
    // (the next pose's staging starts with a barrier, after these reads)
  }
}

template <class P>
__global__ __launch_bounds__(64 * AtomRows<P::N>::XW) __attribute__((amdgpu_waves_per_eu(1, AtomRows<P::N>::XOCC)))
void k_atoms_xz(const RotMeta* __restrict__ meta,
                                                                      const float2* __restrict__ R,
                                                                      unsigned long long* __restrict__ keys,
                                                                      const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  static_assert(N == 96 || N == 128, "the two-kernel atom form: grid edges 96 and 128");
  using AR = AtomRows<N>;
  using RS = RegShape<N>;
  constexpr int L = RS::L, V = RS::V, LPW = RS::LPW;
  constexpr int TS = AR::TS, XRW = AR::XRW, TURNS = (AR::XW + AR::XRW - 1) / AR::XRW;
  extern __shared__ float2 tb[];  // [ZH][TS] (kz, x'), or the transforms' scratch
  __shared__ float2 stwr[L * V];
  const int y = blockIdx.x, s = blockIdx.y;
  const RotMeta m = meta[s];
  if (!m.valid || y >= m.win[1]) return;
  const int wx = m.win[0], wz = m.win[2];
  const int tid = threadIdx.x;
  for (int t = tid; t < L * V; t += blockDim.x) stwr[t] = twg[(t / V) * (t % V)];
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  int ln = lane % LPW, n2 = lane / LPW;
  asm volatile("" : "+v"(ln), "+v"(n2));
  const float2* twr = stwr + n2 * V;
  // line t = LPW * wave + ln: kz = t, and for t = 0 the kz = 0 and N/2 lines
  // as one; value j at kx = n2 + L j
  const int t = LPW * wave + ln;
  const float2* rr = R + (size_t)s * ((size_t)ZH * N * N) + (size_t)y * (ZH * N) + n2;
  float2 v[V], bv[V];
  {
    // the kz = N/2 row's loads first, so none of them waits on another
    if (t == 0) {
      const float2* b = rr + (N / 2) * N;
#pragma unroll
      for (int j = 0; j < V; ++j) bv[j] = b[L * j];
    }
    const float2* a = rr + t * N;
#pragma unroll
    for (int j = 0; j < V; ++j) v[j] = a[L * j];
    if (t == 0) {
#pragma unroll
      for (int j = 0; j < V; ++j) v[j] = make_float2(v[j].x - bv[j].y, v[j].y + bv[j].x);
    }
  }
  __syncthreads();  // stwr
#pragma unroll
  for (int r = 0; r < TURNS; ++r) {
    if (wave / XRW == r) reg_fft<N, 1>(v, twr, tb + (wave % XRW) * RS::SLICE, ln, n2);
    __syncthreads();  // this turn's scratch reads are done (the [kz][x'] tile overlays the scratch)
  }
  // every x' into [kz][x'] (the columns past the window are read as the
  // partners of the window's last pair); line 0 holds kz = 0 in its real
  // and kz = N/2 in its imaginary parts
#pragma unroll
  for (int j = 0; j < V; ++j) tb[t * TS + n2 + L * j] = v[j];
  if (t == 0) {
#pragma unroll
    for (int j = 0; j < V; ++j) tb[(N / 2) * TS + n2 + L * j] = make_float2(v[j].y, 0.f);
  }
  __syncthreads();
  // z lines: x' pair (2p, 2p + 1), p = t, as one complex line A + iB; value j
  // at kz = n2 + L j over the full range (kz > N/2 from the conjugates).  A
  // wavefront whose pairs all lie past the window skips its transforms.
  const int p = t;
  const bool act = 2 * p < wx, bact = 2 * p + 1 < wx;
  const bool wact = 2 * LPW * wave < wx;
  float2 z[V];
#pragma unroll
  for (int j = 0; j < V; ++j) {
    const int k = n2 + L * j;
    const int kk = k <= N / 2 ? k : N - k;
    float2 A = tb[kk * TS + 2 * p], B = tb[kk * TS + 2 * p + 1];
    if (k == 0 || k == N / 2) {
      A.y = 0.f;
      B.y = 0.f;
    } else if (k > N / 2) {
      A.y = -A.y;
      B.y = -B.y;
    }
    z[j] = make_float2(A.x - B.y, A.y + B.x);
  }
  __syncthreads();  // the tile is free: the transforms' scratch
#pragma unroll
  for (int r = 0; r < TURNS; ++r) {
    if (wact && wave / XRW == r) reg_fft<N, 1>(z, twr, tb + (wave % XRW) * RS::SLICE, ln, n2);
    if (r + 1 < TURNS) __syncthreads();
  }
  // the lane's window minimum: values in increasing translation index (line
  // 2p's z values, then line 2p + 1's); the first strict minimum of the
  // ordered energy bits (the keys' high word) and its position, then one key
  unsigned bu = 0xffffffffu;
  int bq = -1;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const bool lok = h ? bact : act;
#pragma unroll
    for (int q = 0; q < V; ++q) {
      const unsigned u = energy_bits(h ? z[q].y : z[q].x);
      const bool take = lok && n2 + L * q < wz && (u < bu || bq < 0);
      bu = take ? u : bu;
      bq = take ? h * V + q : bq;
    }
  }
  unsigned long long best = ~0ull;
  if (bq >= 0) {
    const int h = bq >= V ? 1 : 0, q = bq - h * V;
    best = ((unsigned long long)bu << 32) | (unsigned)(((2 * p + h) * N + y) * N + n2 + L * q);
  }
  block_min_commit<AR::XW>(best, keys + s);
}

size_t atoms_xz_lds_bytes(int N) {
  return (size_t)(N == 96 ? AtomRows<96>::XU : N == 128 ? AtomRows<128>::XU : 0) * sizeof(float2);
}

// k_atoms_xz's static LDS: the twiddles (N float2) and the key reduction.
static size_t atoms_xz_static_lds(int N) {
  return (size_t)N * sizeof(float2) + 16 * sizeof(unsigned long long);
}

// Column form (and the atom form at 64), last stage: one workgroup per
// (pose, block of 64/L * 4 row pairs).  Row pair (x', y0 = L*m + 2h, y0 + 1)
// is one complex line A + iB of the z C2R (two-for-one, imaginary parts of
// the kz = 0 and N/2 bins dropped as a C2R transform does); the window
// minimum of the two real rows goes to the pose's key.
template <class P>
__global__ __launch_bounds__(kBlock) void k_inv_z(const RotMeta* __restrict__ meta, const float2* __restrict__ T2,
                                                  unsigned long long* __restrict__ keys,
                                                  const float2* __restrict__ twg) {
  constexpr int N = P::N, ZH = P::ZH;
  using RS = RegShape<N>;
  constexpr int L = RS::L, V = RS::V, LPW = RS::LPW;
  static_assert(N == 64, "column form and 64-point atom form");
  constexpr int H = L / 2;        // row pairs per L rows
  constexpr int LPB = 4 * LPW;    // row pairs per workgroup
  __shared__ float2 sbuf[4][RS::SLICE];
  const int s = blockIdx.y;
  const RotMeta m = meta[s];
  if (!m.valid) return;
  const int wx = m.win[0], wy = m.win[1], wz = m.win[2];
  const int M = (wy + L - 1) / L;
  const int nl = M * wx * H;
  if ((int)blockIdx.x * LPB >= nl) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ln = lane % LPW, n2 = lane / LPW;
  float2 twr[V];
  reg_twiddles<N>(twg, n2, twr);
  const int l = blockIdx.x * LPB + wave * LPW + ln;
  const bool ok = l < nl;
  const int mm = l / (wx * H), r = l - mm * (wx * H);
  const int xp = r / H, h = r - (r / H) * H;
  const int y0 = L * mm + 2 * h;
  const bool aok = ok && y0 < wy, bok = ok && y0 + 1 < wy;
  const float2* tb = T2 + (size_t)s * ZH * (N * N) + (size_t)(mm * N + xp) * L + 2 * h;
  float2 z[V];
#pragma unroll
  for (int n1 = 0; n1 < V; ++n1) {
    const int k = n2 + L * n1;
    const int kk = k <= N / 2 ? k : N - k;
    float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ok) f = *reinterpret_cast<const float4*>(tb + (size_t)kk * (N * N));
    float2 A = aok ? make_float2(f.x, f.y) : make_float2(0.f, 0.f);
    float2 B = bok ? make_float2(f.z, f.w) : make_float2(0.f, 0.f);
    if (k == 0 || k == N / 2) {
      A.y = 0.f;
      B.y = 0.f;
    } else if (k > N / 2) {
      A.y = -A.y;
      B.y = -B.y;
    }
    z[n1] = make_float2(A.x - B.y, A.y + B.x);
  }
  reg_fft<N, 1>(z, twr, sbuf[wave], ln, n2);
  unsigned long long best = ~0ull;
  if (ok) {
#pragma unroll
    for (int q = 0; q < V; ++q) {
      const int zz = n2 + L * q;
      if (zz < wz) {
        const unsigned lin0 = (unsigned)((xp * N + y0) * N + zz);
        if (aok) {
          const unsigned long long k0 = energy_key(z[q].x, lin0);
          best = k0 < best ? k0 : best;
        }
        if (bok) {
          const unsigned long long k1 = energy_key(z[q].y, lin0 + (unsigned)N);
          best = k1 < best ? k1 : best;
        }
      }
    }
  }
  block_min_commit(best, keys + s);
}

// Rotate.cpp:4-27 for every (quaternion, atom): out[q][a][3].
__global__ __launch_bounds__(kBlock) void k_rotate(int nq, int natom, const float* __restrict__ quats,
                                                   const float* __restrict__ coords, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq * natom) return;
  const int q = i / natom, a = i - (i / natom) * natom;
  float qq[4] = {quats[q * 4 + 0], quats[q * 4 + 1], quats[q * 4 + 2], quats[q * 4 + 3]};
  float o[3];
  rotate_exact(qq, coords[a * 3 + 0], coords[a * 3 + 1], coords[a * 3 + 2], o);
  out[(size_t)i * 3 + 0] = o[0];
  out[(size_t)i * 3 + 1] = o[1];
  out[(size_t)i * 3 + 2] = o[2];
}

// ---------------------------------------------------------------------------
// FillLigandGrid.cpp:7-120 as a standalone device operation on caller-given
// rotated coordinates and minimum corners: placement per (pose, atom), then a
// per-cell gather that adds each channel's atoms in ascending order on top of
// the existing cell value (the reference accumulates with +=).  Grid layout
// [pose][channel][X][Y][Z], channel index pose*C + c (FillLigandGrid.cpp:38).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_place(int nq, int natom, const float* __restrict__ coors,
                                                  const float* __restrict__ mincoors, float h,
                                                  AtomGrid* __restrict__ ag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq * natom) return;
  const int q = i / natom;
  AtomGrid g;
  grid_place(coors[(size_t)i * 3 + 0], mincoors[q * 3 + 0], h, g.ix, g.fx);
  grid_place(coors[(size_t)i * 3 + 1], mincoors[q * 3 + 1], h, g.iy, g.fy);
  grid_place(coors[(size_t)i * 3 + 2], mincoors[q * 3 + 2], h, g.iz, g.fz);
  g.pad = 0;
  ag[i] = g;
}

__global__ __launch_bounds__(kBlock) void k_fill_generic(int natom, const AtomGrid* __restrict__ ag,
                                                         const int* __restrict__ ch_start,
                                                         const int* __restrict__ chan_atoms,
                                                         const float* __restrict__ wv, const float* __restrict__ wq,
                                                         int C, int X, int Y, int Z, float* __restrict__ grid) {
  __shared__ SpreadAtom sa[kMaxAtoms];
  const int q = blockIdx.z, c = blockIdx.y;
  const size_t G = (size_t)X * Y * Z;
  const int b0 = ch_start[c], nc = ch_start[c + 1] - ch_start[c];
  for (int i = threadIdx.x; i < nc; i += blockDim.x) {
    const int a = chan_atoms[b0 + i];
    const AtomGrid g = ag[(size_t)q * natom + a];
    SpreadAtom t;
    t.ix = g.ix;
    t.iy = g.iy;
    t.iz = g.iz;
    t.w = (c == C - 1) ? wq[a] : wv[a];
    t.fx = g.fx;
    t.fy = g.fy;
    t.fz = g.fz;
    sa[i] = t;
  }
  __syncthreads();
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= G) return;
  const int x = (int)(p / ((size_t)Y * Z));
  const int y = (int)((p / Z) % Y);
  const int z = (int)(p % Z);
  float* cell = grid + ((size_t)q * C + c) * G + p;
  float acc = *cell;
  bool touched = false;
  for (int i = 0; i < nc; ++i) {
    const int dx = x - sa[i].ix, dy = y - sa[i].iy, dz = z - sa[i].iz;
    if ((unsigned)dx <= 1u && (unsigned)dy <= 1u && (unsigned)dz <= 1u) {
      acc = accum_corner(acc, x_factor(sa[i].w, sa[i].fx, dx), sa[i].fy, sa[i].fz, dy, dz);
      touched = true;
    }
  }
  if (touched) *cell = acc;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

#define CDK_DISPATCH_N(N, ...)                         \
  switch (N) {                                         \
    case 32: { using PL = P32a; __VA_ARGS__ } break;   \
    case 48: { using PL = P48a; __VA_ARGS__ } break;   \
    case 64: { using PL = P64a; __VA_ARGS__ } break;   \
    case 96: { using PL = P96a; __VA_ARGS__ } break;   \
    case 128: { using PL = P128a; __VA_ARGS__ } break; \
    default: return hipErrorInvalidValue;              \
  }

bool grid_edge_supported(int N) { return N == 32 || N == 48 || N == 64 || N == 96 || N == 128; }

hipError_t launch_prep(hipStream_t st, const PoseSlot* slots, int nslots, const LigandDesc* ligs,
                       const float* coords, int N, float h, int astride, int boxmax, float* rot, AtomGrid* ag,
                       RotMeta* meta, unsigned long long* keys) {
  if (nslots <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_prep, dim3(nslots), dim3(kPrepBlock), 0, st, slots, nslots, ligs, coords, N, h, astride, boxmax,
                     rot, ag, meta, keys);
  return hipGetLastError();
}

hipError_t launch_spread_full(hipStream_t st, const PoseSlot* slots, int nslots, const LigandDesc* ligs,
                              const int* chan_atoms, const float* wv, const float* wq, const AtomGrid* ag,
                              int astride, const RotMeta* meta, int N, int C, float* out) {
  if (nslots <= 0) return hipSuccess;
  const size_t G = (size_t)N * N * N;
  hipLaunchKernelGGL(k_spread_full, dim3(ceil_div(G, kBlock), C, nslots), dim3(kBlock), 0, st, slots, ligs,
                     chan_atoms, wv, wq, ag, astride, meta, N, C, out);
  return hipGetLastError();
}

hipError_t launch_r2c_3d(hipStream_t st, int N, const float* in, float2* out, int nb, const float2* tw) {
  if (nb <= 0) return hipSuccess;
  CDK_DISPATCH_N(N, {
    hipLaunchKernelGGL(k_r2c_planes<PL>, dim3(nb * N), dim3(kBlock), 0, st, in, out, tw);
    hipLaunchKernelGGL((k_xlines<PL, -1>), dim3(nb * N), dim3(kBlock), 0, st, out, tw);
  });
  return hipGetLastError();
}

hipError_t launch_c2r_3d(hipStream_t st, int N, float2* in, float* out, int nb, const float2* tw) {
  if (nb <= 0) return hipSuccess;
  CDK_DISPATCH_N(N, {
    hipLaunchKernelGGL((k_xlines<PL, 1>), dim3(nb * N), dim3(kBlock), 0, st, in, tw);
    hipLaunchKernelGGL(k_c2r_planes<PL>, dim3(nb * N), dim3(kBlock), 0, st, in, out, tw);
  });
  return hipGetLastError();
}

hipError_t launch_conj_mult(hipStream_t st, float2* L, const float2* P, const int* pmap, int C, long long H,
                            int nb) {
  const long long total = (long long)nb * C * H;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_conj_mult, dim3(ceil_div(total, kBlock)), dim3(kBlock), 0, st, L, P, pmap, C, H, total);
  return hipGetLastError();
}

hipError_t launch_sum_grids(hipStream_t st, const float2* L, float2* S, int C, long long H, int nb, float G) {
  const long long total = (long long)nb * H;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sum_grids, dim3(ceil_div(total, kBlock)), dim3(kBlock), 0, st, L, S, C, H, total, G);
  return hipGetLastError();
}

hipError_t launch_argmin_full(hipStream_t st, const float* E, const RotMeta* meta, unsigned long long* keys,
                              int N, int nb) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_argmin_full, dim3(N, nb), dim3(kBlock), 0, st, E, meta, keys, N);
  return hipGetLastError();
}

size_t fwd_reg_lds_bytes(int N, int boxmax, int max_chan_atoms) {
  int RSTR = 0;
  switch (N) {
    case 32: RSTR = FwdRegShape<32>::RSTR; break;
    case 48: RSTR = FwdRegShape<48>::RSTR; break;
    case 64: RSTR = FwdRegShape<64>::RSTR; break;
    case 96: RSTR = FwdRegShape<96>::RSTR; break;
    case 128: RSTR = FwdRegShape<128>::RSTR; break;
    default: return 0;
  }
  const int npmax = (boxmax + 1) / 2;
  return (size_t)std::max(1, max_chan_atoms) * sizeof(SpreadAtom) + (size_t)npmax * RSTR * sizeof(float2);
}

template <int N>
constexpr size_t fwd_reg_static() {
  using R = RegShape<N>;
  return (4 * (size_t)R::SLICE + (R::V >= 16 ? (size_t)R::L * R::V : 1)) * sizeof(float2);
}

namespace {
size_t g_max_dyn_lds = 64 * 1024;
}

hipError_t configure_fwd_reg(size_t max_dyn_lds) {
  g_max_dyn_lds = max_dyn_lds;
  const int sizes[] = {32, 48, 64, 96, 128};
  for (int N : sizes) {
    hipError_t e = hipSuccess;
    CDK_DISPATCH_N(N, {
      // the dynamic part plus the static arrays must fit the device
      const size_t stat = fwd_reg_static<PL::N>();
      const size_t dyn = max_dyn_lds > stat ? max_dyn_lds - stat : 0;
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fwd_planes_reg<PL>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
      if constexpr (PL::N == 64) {
        if (e == hipSuccess)
          e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fwd_z<PL>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
      }
      if constexpr (PL::N == 96 || PL::N == 128) {
        if (e == hipSuccess && atoms_xz_lds_bytes(PL::N) + atoms_xz_static_lds(PL::N) <= max_dyn_lds)
          e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_atoms_xz<PL>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)atoms_xz_lds_bytes(PL::N));
      }
    });
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

static size_t fwd_reg_static_lds(int N) {
  switch (N) {
    case 32: return fwd_reg_static<32>();
    case 48: return fwd_reg_static<48>();
    case 64: return fwd_reg_static<64>();
    case 96: return fwd_reg_static<96>();
    case 128: return fwd_reg_static<128>();
    default: return 0;
  }
}

bool fwd_reg_supported(int N) {
  return fwd_reg_lds_bytes(N, N, kMaxAtoms) + fwd_reg_static_lds(N) <= g_max_dyn_lds;
}

bool col_supported(int N) { return N == 64 && fwd_reg_supported(N); }

hipError_t launch_fwd_z(hipStream_t st, int N, const PoseSlot* slots, int nslots, const LigandDesc* ligs,
                        const int* chan_atoms, const float* wv, const float* wq, const AtomGrid* ag, int astride,
                        const RotMeta* meta, int cmax, int boxmax, int max_chan_atoms, float2* Z, const float2* tw) {
  if (nslots <= 0) return hipSuccess;
  if (!col_supported(N)) return hipErrorInvalidValue;
  const int natom = std::max(1, max_chan_atoms);
  const size_t dyn = fwd_reg_lds_bytes(N, boxmax, natom);
  if (dyn == 0 || dyn + fwd_reg_static_lds(N) > g_max_dyn_lds) return hipErrorInvalidValue;
  CDK_DISPATCH_N(N, {
    if constexpr (PL::N == 64)
      hipLaunchKernelGGL(k_fwd_z<PL>, dim3(cmax, nslots), dim3(kBlock), dyn, st, slots, ligs, chan_atoms, wv, wq, ag,
                         astride, meta, cmax, boxmax, natom, Z, tw);
  });
  return hipGetLastError();
}

hipError_t launch_mid_kz(hipStream_t st, int N, const PoseSlot* slots, const LigandDesc* ligs, const RotMeta* meta,
                         int nslots, int group, const float2* Z, int cmax, int boxmax, const float2* P2, float2* T2,
                         float invG, const float2* tw) {
  if (nslots <= 0) return hipSuccess;
  if (!col_supported(N) || group <= 0) return hipErrorInvalidValue;
  const dim3 grid(xcd_unit_grid(N / 16, 1, (int)ceil_div(nslots, group)));
  CDK_DISPATCH_N(N, {
    if constexpr (PL::N == 64) {
      // poses whose box fits in N/2 along x go through the half-plane form,
      // the others (if the launch can hold any) through the full form
      hipLaunchKernelGGL((k_mid_kz<PL, true>), grid, dim3(kBlock), 0, st, slots, ligs, meta, nslots, group, Z, cmax,
                         boxmax, P2, T2, invG, tw, 1);
      if (2 * boxmax > N)
        hipLaunchKernelGGL((k_mid_kz<PL, false>), grid, dim3(kBlock), 0, st, slots, ligs, meta, nslots, group, Z,
                           cmax, boxmax, P2, T2, invG, tw, 1);
    }
  });
  return hipGetLastError();
}

hipError_t launch_inv_z(hipStream_t st, int N, const RotMeta* meta, int nslots, const float2* T2,
                        unsigned long long* keys, const float2* tw) {
  if (nslots <= 0) return hipSuccess;
  if (!col_supported(N) && !atoms_supported(N)) return hipErrorInvalidValue;
  CDK_DISPATCH_N(N, {
    // windows are at most N - 2 wide (boxes are at least 3 cells)
    if constexpr (PL::N == 64) {
      using RS = RegShape<PL::N>;
      const int w = PL::N - 2;
      const int pairs = ((w + RS::L - 1) / RS::L) * w * (RS::L / 2);
      const unsigned gx = ceil_div(pairs, 4 * RS::LPW);
      hipLaunchKernelGGL(k_inv_z<PL>, dim3(gx, nslots), dim3(kBlock), 0, st, meta, T2, keys, tw);
    }
    if constexpr (PL::N == 96 || PL::N == 128)
      hipLaunchKernelGGL(k_atoms_xz<PL>, dim3(PL::N - 2, nslots), dim3(64 * AtomRows<PL::N>::XW),
                         atoms_xz_lds_bytes(PL::N), st, meta, T2, keys, tw);
  });
  return hipGetLastError();
}

bool atoms_supported(int N) {
  // the x-z kernel's tile or scratch (dynamic) and its twiddles and key
  // reduction (static)
  return N == 64 || ((N == 96 || N == 128) && atoms_xz_lds_bytes(N) + atoms_xz_static_lds(N) <= g_max_dyn_lds);
}

hipError_t launch_mid_atoms(hipStream_t st, int N, const PoseSlot* slots, const LigandDesc* ligs,
                            const int* pad_atoms, const float* wv, const float* wq, const AtomGrid* ag, int astride,
                            const RotMeta* meta, int nslots, int group, const float2* P3, float2* T2, float invG,
                            const float2* tw, int products) {
  if (nslots <= 0) return hipSuccess;
  if (!atoms_supported(N) || group <= 0 || (products != kAtomF16 && products != kAtomF32 && products != kAtomF16x3))
    return hipErrorInvalidValue;
  const dim3 grid(xcd_plane_grid((N / 2 + 1) / 8, (N / 2 + 1) % 8, (int)ceil_div(nslots, group)));
  CDK_DISPATCH_N(N, {
    if constexpr (PL::N == 64)
      hipLaunchKernelGGL((products == kAtomF16x3 ? k_mid_atoms<PL, kAtomF16x3>
                          : products == kAtomF16 ? k_mid_atoms<PL, kAtomF16>
                                                 : k_mid_atoms<PL, kAtomF32>),
                         grid, dim3(kBlock), 0, st, slots, ligs, pad_atoms, wv, wq, ag, astride, meta, nslots, group,
                         P3, T2, invG, tw);
    if constexpr (PL::N == 96 || PL::N == 128) {
      constexpr int NB = PL::N / 32;
      const dim3 gf(xcd_plane_grid(PL::ZH / 8, PL::ZH % 8, NB * (int)ceil_div(nslots, group)));
      hipLaunchKernelGGL((products == kAtomF16x3 ? k_atoms_fwd<PL, kAtomF16x3>
                          : products == kAtomF16 ? k_atoms_fwd<PL, kAtomF16>
                                                 : k_atoms_fwd<PL, kAtomF32>),
                         gf, dim3(64 * NB), 0, st, slots, ligs, pad_atoms, wv, wq, ag, astride, meta, nslots, group,
                         P3, T2, invG, tw);
    }
  });
  return hipGetLastError();
}

hipError_t launch_to_mfma_blocks(hipStream_t st, int N, const float2* in, float2* out, int nch) {
  const size_t n = (size_t)nch * N * N * (N / 2 + 1);
  if (n == 0) return hipSuccess;
  if (!atoms_supported(N)) return hipErrorInvalidValue;
  CDK_DISPATCH_N(N, {
    if constexpr (PL::N == 64 || PL::N == 96 || PL::N == 128)
      hipLaunchKernelGGL(k_to_mfma_blocks<PL>, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, st, in, out, nch);
  });
  return hipGetLastError();
}

hipError_t launch_to_kz_blocks(hipStream_t st, int N, const float2* in, float2* out, int nch) {
  const size_t n = (size_t)nch * N * N * (N / 2 + 1);
  if (n == 0) return hipSuccess;
  if (N != 64) return hipErrorInvalidValue;
  CDK_DISPATCH_N(N, {
    if constexpr (PL::N == 64)
      hipLaunchKernelGGL(k_to_kz_blocks<PL>, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, st, in, out, nch);
  });
  return hipGetLastError();
}

hipError_t launch_fwd_planes(hipStream_t st, int N, int variant, const PoseSlot* slots, int nslots,
                             const LigandDesc* ligs, const int* chan_atoms, const float* wv, const float* wq,
                             const AtomGrid* ag, int astride, const RotMeta* meta, int cmax, int boxmax,
                             int max_chan_atoms, float2* A, const float2* tw) {
  if (nslots <= 0) return hipSuccess;
  const int natom = std::max(1, max_chan_atoms);
  if (variant == kFwdReg) {
    const size_t dyn = fwd_reg_lds_bytes(N, boxmax, natom);
    if (dyn == 0 || dyn + fwd_reg_static_lds(N) > g_max_dyn_lds) return hipErrorInvalidValue;
    const int npmax = (boxmax + 1) / 2;
    CDK_DISPATCH_N(N, {
      hipLaunchKernelGGL(k_fwd_planes_reg<PL>, dim3(cmax, nslots), dim3(kBlock), dyn, st, slots, ligs, chan_atoms,
                         wv, wq, ag, astride, meta, cmax, boxmax, natom, npmax, A, tw);
    });
    return hipGetLastError();
  }
  const size_t dyn = (size_t)natom * sizeof(SpreadAtom);
  CDK_DISPATCH_N(N, {
    hipLaunchKernelGGL(k_fwd_planes<PL>, dim3(cmax, nslots), dim3(kBlock), dyn, st, slots, ligs, chan_atoms, wv, wq,
                       ag, astride, meta, cmax, boxmax, A, tw);
  });
  return hipGetLastError();
}

hipError_t launch_to_line_blocks(hipStream_t st, int N, const float2* in, float2* out, int nch) {
  const size_t n = (size_t)nch * N * N * (N / 2 + 1);
  if (n == 0) return hipSuccess;
  CDK_DISPATCH_N(N, {
    hipLaunchKernelGGL(k_to_line_blocks<PL>, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, st, in, out, nch);
  });
  return hipGetLastError();
}

int mid_line_groups(int N) {
  const int L = (N == 64 || N == 128) ? 8 : 4;
  const int lpb = 4 * (64 / L);
  const int NL = N * (N / 2 + 1);
  return (NL + lpb - 1) / lpb;
}

namespace {
template <class PL, int D, bool H>
void mid_launch_one(dim3 grid, hipStream_t st, const PoseSlot* slots, const LigandDesc* ligs, const RotMeta* meta,
                    int nslots, int group, const float2* A, int cmax, int boxmax, const float2* Phat, float2* T,
                    float invG, const float2* tw) {
  hipLaunchKernelGGL((k_mid<PL, D, H>), grid, dim3(kBlock), 0, st, slots, ligs, meta, nslots, group, A, cmax, boxmax,
                     Phat, T, invG, tw);
}
template <class PL, int NP, bool H>
void mid_np_launch_one(dim3 grid, hipStream_t st, const PoseSlot* slots, const LigandDesc* ligs, const RotMeta* meta,
                       int nslots, int group, const float2* A, int cmax, int boxmax, const float2* Phat, float2* T,
                       float invG, const float2* tw) {
  hipLaunchKernelGGL((k_mid_np<PL, NP, H>), grid, dim3(kBlock), 0, st, slots, ligs, meta, nslots, group, A, cmax,
                     boxmax, Phat, T, invG, tw);
}
}  // namespace

int mid_max_depth(int N) { return (N == 32 || N == 64) ? 3 : 1; }
int mid_max_poses(int N) { return (N == 32 || N == 64) ? 4 : 1; }

hipError_t launch_mid(hipStream_t st, int N, int depth, int np, const PoseSlot* slots, const LigandDesc* ligs,
                      const RotMeta* meta, int nslots, int group, const float2* A, int cmax, int boxmax,
                      const float2* Phat, int npch, float2* T, float invG, const float2* tw) {
  (void)npch;
  if (nslots <= 0) return hipSuccess;
  if (group <= 0 || np < 1 || np > mid_max_poses(N) || (np & (np - 1)) != 0) return hipErrorInvalidValue;
  const int nlg = mid_line_groups(N);
  const dim3 grid(xcd_unit_grid(nlg / 8, nlg % 8, (int)ceil_div(nslots, group)));
  const bool half = 2 * boxmax <= N;  // every valid pose's box fits in N/2
  if (np > 1) {
    CDK_DISPATCH_N(N, {
      if constexpr (RegShape<PL::N>::V <= 8) {
        if (np == 4) {
          if (half)
            mid_np_launch_one<PL, 4, true>(grid, st, slots, ligs, meta, nslots, group, A, cmax, boxmax, Phat, T, invG, tw);
          else
            mid_np_launch_one<PL, 4, false>(grid, st, slots, ligs, meta, nslots, group, A, cmax, boxmax, Phat, T, invG, tw);
        } else {
          if (half)
            mid_np_launch_one<PL, 2, true>(grid, st, slots, ligs, meta, nslots, group, A, cmax, boxmax, Phat, T, invG, tw);
          else
            mid_np_launch_one<PL, 2, false>(grid, st, slots, ligs, meta, nslots, group, A, cmax, boxmax, Phat, T, invG, tw);
        }
      }
    });
    return hipGetLastError();
  }
  CDK_DISPATCH_N(N, {
    if constexpr (RegShape<PL::N>::V <= 8) {
      if (depth >= 3) {
        if (half)
          mid_launch_one<PL, 3, true>(grid, st, slots, ligs, meta, nslots, group, A, cmax, boxmax, Phat, T, invG, tw);
        else
          mid_launch_one<PL, 3, false>(grid, st, slots, ligs, meta, nslots, group, A, cmax, boxmax, Phat, T, invG, tw);
      } else {
        if (half)
          mid_launch_one<PL, 2, true>(grid, st, slots, ligs, meta, nslots, group, A, cmax, boxmax, Phat, T, invG, tw);
        else
          mid_launch_one<PL, 2, false>(grid, st, slots, ligs, meta, nslots, group, A, cmax, boxmax, Phat, T, invG, tw);
      }
    } else {
      if (half)
        mid_launch_one<PL, 1, true>(grid, st, slots, ligs, meta, nslots, group, A, cmax, boxmax, Phat, T, invG, tw);
      else
        mid_launch_one<PL, 1, false>(grid, st, slots, ligs, meta, nslots, group, A, cmax, boxmax, Phat, T, invG, tw);
    }
  });
  return hipGetLastError();
}

bool inv_reg_supported(int N) { return grid_edge_supported(N); }

hipError_t launch_inv_argmin(hipStream_t st, int N, int variant, const RotMeta* meta, int nslots, int wmax,
                             const float2* T, unsigned long long* keys, const float2* tw) {
  if (nslots <= 0) return hipSuccess;
  if (variant == kInvReg) {
    if (!inv_reg_supported(N)) return hipErrorInvalidValue;
    CDK_DISPATCH_N(N, {
      const int gx = (wmax + InvRegShape<PL::N>::XPB - 1) / InvRegShape<PL::N>::XPB;
      hipLaunchKernelGGL(k_inv_argmin_reg<PL>, dim3(gx, nslots), dim3(kBlock), 0, st, meta, T, keys, tw);
    });
    return hipGetLastError();
  }
  CDK_DISPATCH_N(N, {
    hipLaunchKernelGGL(k_inv_argmin<PL>, dim3(wmax, nslots), dim3(kBlock), 0, st, meta, T, keys, tw);
  });
  return hipGetLastError();
}

hipError_t launch_rotate(hipStream_t st, int nq, int natom, const float* quats, const float* coords, float* out) {
  const size_t n = (size_t)nq * natom;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rotate, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, st, nq, natom, quats, coords, out);
  return hipGetLastError();
}

hipError_t launch_place(hipStream_t st, int nq, int natom, const float* coors, const float* mincoors, float h,
                        AtomGrid* ag) {
  const size_t n = (size_t)nq * natom;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_place, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, st, nq, natom, coors, mincoors, h, ag);
  return hipGetLastError();
}

hipError_t launch_fill_generic(hipStream_t st, int nq, int natom, const AtomGrid* ag, const int* ch_start,
                               const int* chan_atoms, const float* wv, const float* wq, int C, int X, int Y, int Z,
                               float* grid) {
  const size_t G = (size_t)X * Y * Z;
  if (nq <= 0 || G == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_generic, dim3(ceil_div(G, kBlock), C, nq), dim3(kBlock), 0, st, natom, ag, ch_start,
                     chan_atoms, wv, wq, C, X, Y, Z, grid);
  return hipGetLastError();
}

}  // namespace cdk
