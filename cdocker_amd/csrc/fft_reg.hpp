// Register-resident four-step FFT of N = V * L points run by one wavefront.
//
// A line is held by L lanes, V values each: lane (line, n2) holds
// x[n2 + L*n1] in register n1.  The transform is
//   1. DFT_V over the lane's registers,
//   2. twiddle W_N^(n2*k1),
//   3. one transpose through LDS,
//   4. V/L DFT_L per lane over the values gathered from the L lanes,
// and leaves lane (line, j) holding X[j + L*m] in register m -- the same
// layout as the input, so a forward transform, a pointwise product and an
// inverse transform chain with no reordering.  64/L lines per wavefront.
//
// The transpose buffer holds V groups (one per k1) of L rows x LPW lines,
// each group padded to K = L*LPW + PAD float2.  Lane (line, n2) writes
// value k1 at k1*K + n2*LPW + line and reads row m2 of group j + L*q at
// (j + L*q)*K + m2*LPW + line, so every address is a per-lane base plus a
// compile-time offset (one address register each way).  PAD = 8 (L = 8) or
// 16 (L = 4) shifts consecutive groups by half a bank row: the 16 or 32
// lanes served together then hit distinct 8-byte bank slots for the writes
// (consecutive lines of one group) and for the reads (one row from each of
// the lanes' groups).
#pragma once
#include "fft_lds.hpp"

namespace cdk {

template <int M> struct RootTable;

template <> struct RootTable<12> {
  static __device__ __forceinline__ float cosv(int m) {
    constexpr float t[12] = {1.000000000e+00f, 8.660254038e-01f, 5.000000000e-01f, 0.0f, -5.000000000e-01f, -8.660254038e-01f, -1.000000000e+00f, -8.660254038e-01f, -5.000000000e-01f, 0.0f, 5.000000000e-01f, 8.660254038e-01f};
    return t[m];
  }
  static __device__ __forceinline__ float sinv(int m) {
    constexpr float t[12] = {0.000000000e+00f, 5.000000000e-01f, 8.660254038e-01f, 1.000000000e+00f, 8.660254038e-01f, 5.000000000e-01f, 0.0f, -5.000000000e-01f, -8.660254038e-01f, -1.000000000e+00f, -8.660254038e-01f, -5.000000000e-01f};
    return t[m];
  }
};

template <> struct RootTable<24> {
  static __device__ __forceinline__ float cosv(int m) {
    constexpr float t[24] = {1.000000000e+00f, 9.659258263e-01f, 8.660254038e-01f, 7.071067812e-01f, 5.000000000e-01f, 2.588190451e-01f, 0.0f, -2.588190451e-01f, -5.000000000e-01f, -7.071067812e-01f, -8.660254038e-01f, -9.659258263e-01f, -1.000000000e+00f, -9.659258263e-01f, -8.660254038e-01f, -7.071067812e-01f, -5.000000000e-01f, -2.588190451e-01f, 0.0f, 2.588190451e-01f, 5.000000000e-01f, 7.071067812e-01f, 8.660254038e-01f, 9.659258263e-01f};
    return t[m];
  }
  static __device__ __forceinline__ float sinv(int m) {
    constexpr float t[24] = {0.000000000e+00f, 2.588190451e-01f, 5.000000000e-01f, 7.071067812e-01f, 8.660254038e-01f, 9.659258263e-01f, 1.000000000e+00f, 9.659258263e-01f, 8.660254038e-01f, 7.071067812e-01f, 5.000000000e-01f, 2.588190451e-01f, 0.0f, -2.588190451e-01f, -5.000000000e-01f, -7.071067812e-01f, -8.660254038e-01f, -9.659258263e-01f, -1.000000000e+00f, -9.659258263e-01f, -8.660254038e-01f, -7.071067812e-01f, -5.000000000e-01f, -2.588190451e-01f};
    return t[m];
  }
};

// Composite small DFT R1*R2 in registers (Cooley-Tukey, input index
// R2*n1 + n2, output index k1 + R1*k2).
template <int R1, int R2, int S>
__device__ __forceinline__ void compose_dft(float2* v) {
  constexpr int M = R1 * R2;
  float2 z[R2][R1];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) z[n2][n1] = v[R2 * n1 + n2];
    SmallDFT<R1, S>::run(z[n2]);
  }
#pragma unroll
  for (int n2 = 1; n2 < R2; ++n2)
#pragma unroll
    for (int k1 = 1; k1 < R1; ++k1) {
      const int m = (n2 * k1) % M;
      const float2 w = make_float2(RootTable<M>::cosv(m), S > 0 ? RootTable<M>::sinv(m) : -RootTable<M>::sinv(m));
      z[n2][k1] = c_mul(z[n2][k1], w);
    }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    float2 t[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) t[n2] = z[n2][k1];
    SmallDFT<R2, S>::run(t);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = t[k2];
  }
}

template <int R, int S> struct RegDFT {
  static __device__ __forceinline__ void run(float2* v) { SmallDFT<R, S>::run(v); }
};
template <int S> struct RegDFT<12, S> {
  static __device__ __forceinline__ void run(float2* v) { compose_dft<3, 4, S>(v); }
};
template <int S> struct RegDFT<24, S> {
  static __device__ __forceinline__ void run(float2* v) { compose_dft<3, 8, S>(v); }
};

// DFT of R points whose upper R/2 inputs are zero (the lower half alone is
// read).  The first radix-2 split of such a transform is a copy, which saves
// a quarter of the adds of DFT_8 and DFT_16.  Lengths without a split here
// clear the upper half and run the full transform.
template <int R, int S> struct RegDFTHalf {
  static __device__ __forceinline__ void run(float2* v) {
#pragma unroll
    for (int i = R / 2; i < R; ++i) v[i] = make_float2(0.f, 0.f);
    RegDFT<R, S>::run(v);
  }
};

template <int S> struct RegDFTHalf<8, S> {
  static __device__ __forceinline__ void run(float2* v) {
    const float kR = 0.70710678118654752440f;
    // DFT_4 of (a, b, 0, 0) = (a + b, a + S i b, a - b, a - S i b)
    const float2 ib0 = c_muli<S>(v[2]), ib1 = c_muli<S>(v[3]);
    const float2 e0 = c_add(v[0], v[2]), e1 = c_add(v[0], ib0), e2 = c_sub(v[0], v[2]), e3 = c_sub(v[0], ib0);
    const float2 o0 = c_add(v[1], v[3]), oa = c_add(v[1], ib1), o2r = c_sub(v[1], v[3]), ob = c_sub(v[1], ib1);
    // o1, o3 are the products before the 1/sqrt2 scale
    const float2 o1 = c_add(oa, c_muli<S>(oa));
    const float2 o2 = c_muli<S>(o2r);
    const float2 o3 = c_sub(c_muli<S>(ob), ob);
    v[0] = c_add(e0, o0);
    v[4] = c_sub(e0, o0);
    v[1] = c_fma_s(o1, kR, e1);
    v[5] = c_fnma_s(o1, kR, e1);
    v[2] = c_add(e2, o2);
    v[6] = c_sub(e2, o2);
    v[3] = c_fma_s(o3, kR, e3);
    v[7] = c_fnma_s(o3, kR, e3);
  }
};

template <int S> struct RegDFTHalf<16, S> {
  static __device__ __forceinline__ void run(float2* v) {
    const float C[8] = {1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
                        0.0f, -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f};
    const float Sn[8] = {0.0f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f,
                         1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f};
    float2 e[8], o[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    RegDFTHalf<8, S>::run(e);
    RegDFTHalf<8, S>::run(o);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float2 w = make_float2(C[k], S > 0 ? Sn[k] : -Sn[k]);
      const float2 t = (k == 0) ? o[0] : c_mul(o[k], w);
      v[k] = c_add(e[k], t);
      v[k + 8] = c_sub(e[k], t);
    }
  }
};

// Lanes per line for each transform length.
template <int N> struct RegPlan;
template <> struct RegPlan<32> { static constexpr int L = 4; };
template <> struct RegPlan<48> { static constexpr int L = 4; };
template <> struct RegPlan<64> { static constexpr int L = 8; };
template <> struct RegPlan<96> { static constexpr int L = 4; };
template <> struct RegPlan<128> { static constexpr int L = 8; };

template <int N> struct RegShape {
  static constexpr int L = RegPlan<N>::L;
  static constexpr int V = N / L;
  static constexpr int LPW = 64 / L;  // lines per wavefront
  static constexpr int PAD = L == 8 ? 8 : 16;
  static constexpr int K = L * LPW + PAD;  // float2 per k1 group
  static constexpr int SLICE = V * K;      // float2 per wavefront
  static_assert(V % L == 0, "four-step split needs V divisible by L");
  static_assert(L == 8 || L == 4, "padding chosen for 4 or 8 lanes per line");
};

// Per-lane twiddles W_N^(n2*k1), k1 < V, from the forward table
// tw[m] = exp(-2 pi i m / N).
template <int N>
__device__ __forceinline__ void reg_twiddles(const float2* __restrict__ tw, int n2, float2* twl) {
#pragma unroll
  for (int k1 = 0; k1 < RegShape<N>::V; ++k1) twl[k1] = tw[n2 * k1];
}

// One transform of the wavefront's lines.  `lds` is the wavefront's slice of
// RegShape<N>::SLICE float2.  ln = line within the wavefront, n2 = lane
// within the line.
// HALF: inputs x[n] with n >= N/2 are zero (registers V/2.. are not read).
template <int N, int S, bool HALF = false>
__device__ __forceinline__ void reg_fft(float2* v, const float2* twl, float2* lds, int ln, int n2) {
  constexpr int L = RegShape<N>::L, V = RegShape<N>::V;
  if constexpr (HALF)
    RegDFTHalf<V, S>::run(v);
  else
    RegDFT<V, S>::run(v);
#pragma unroll
  for (int k1 = 1; k1 < V; ++k1) {
    float2 w = twl[k1];
    if (S > 0) w.y = -w.y;
    v[k1] = c_mul(v[k1], w);
  }
  constexpr int K = RegShape<N>::K, LPW = RegShape<N>::LPW;
  WaveSync::sync();  // previous reads of the slice are done
  float2* wb = lds + n2 * LPW + ln;
#pragma unroll
  for (int k1 = 0; k1 < V; ++k1) wb[k1 * K] = v[k1];
  WaveSync::sync();
  const float2* rb = lds + n2 * K + ln;  // group j = n2
#pragma unroll
  for (int q = 0; q < V / L; ++q) {
    float2 u[L];
#pragma unroll
    for (int m2 = 0; m2 < L; ++m2) u[m2] = rb[q * L * K + m2 * LPW];
    SmallDFT<L, S>::run(u);
#pragma unroll
    for (int k2 = 0; k2 < L; ++k2) v[q + (V / L) * k2] = u[k2];
  }
}

}  // namespace cdk
