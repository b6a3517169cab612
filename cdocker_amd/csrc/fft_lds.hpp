// LDS-staged mixed-radix Stockham FFT building blocks for gfx950.
//
// A "line" is a 1-D transform of length N that lives in LDS with an element
// stride `es` and a line stride `ls`.  Every pass reads one buffer and writes
// the other (auto-sort Stockham), so after P passes the result is in natural
// order in buffer (P % 2).  All threads of the workgroup cooperate on all
// lines of a pass; passes are separated by __syncthreads().
//
// Conventions (the same as the unnormalised 3-D R2C / C2R transforms the
// reference's search drives, TranRotaConfSearchNoMini.cu:185-199,222-252):
//   forward  X[k] = sum_n x[n] exp(-2 pi i k n / N)     (SIGN = -1)
//   inverse  x[n] = sum_k X[k] exp(+2 pi i k n / N)     (SIGN = +1, no 1/N)
//
// Twiddles come from a table tw[m] = exp(-2 pi i m / N), m in [0, N), that the
// host computes in double and rounds to float once.
#pragma once
#include <hip/hip_runtime.h>

namespace cdk {

// Every multiply in the transforms is written here, with contraction off and
// the fused multiply-adds spelled out (fmaf).  The compiler then has no
// multiply it may fuse on its own, so a transform rounds the same way in
// every kernel that inlines it, whatever code surrounds it.
__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_scale(float2 a, float s) {
#pragma clang fp contract(off)
  return make_float2(a.x * s, a.y * s);
}
// b + a*s and b - a*s, one rounding each
__device__ __forceinline__ float2 c_fma_s(float2 a, float s, float2 b) {
  return make_float2(fmaf(a.x, s, b.x), fmaf(a.y, s, b.y));
}
__device__ __forceinline__ float2 c_fnma_s(float2 a, float s, float2 b) {
  return make_float2(fmaf(-a.x, s, b.x), fmaf(-a.y, s, b.y));
}
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
#pragma clang fp contract(off)
  return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
// conj(a) * b
__device__ __forceinline__ float2 c_conjmul(float2 a, float2 b) {
#pragma clang fp contract(off)
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.x, b.y, -(a.y * b.x)));
}
// multiply by S*i  (S = +1 or -1)
template <int S>
__device__ __forceinline__ float2 c_muli(float2 a) {
  return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// In-register small DFTs  y_k = sum_n v_n exp(S 2 pi i n k / R)
// ---------------------------------------------------------------------------
template <int R, int S> struct SmallDFT;

template <int S> struct SmallDFT<2, S> {
  static __device__ __forceinline__ void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = c_add(a, b);
    v[1] = c_sub(a, b);
  }
};

template <int S> struct SmallDFT<3, S> {
  static __device__ __forceinline__ void run(float2* v) {
    const float kS3 = 0.86602540378443864676f;  // sqrt(3)/2
    float2 a = v[0];
    float2 t = c_add(v[1], v[2]);
    float2 d = c_sub(v[1], v[2]);
    const float2 m = c_fnma_s(t, 0.5f, a);
    const float2 r = c_muli<S>(d);
    v[0] = c_add(a, t);
    v[1] = c_fma_s(r, kS3, m);
    v[2] = c_fnma_s(r, kS3, m);
  }
};

template <int S> struct SmallDFT<4, S> {
  static __device__ __forceinline__ void run(float2* v) {
    float2 s0 = c_add(v[0], v[2]), d0 = c_sub(v[0], v[2]);
    float2 s1 = c_add(v[1], v[3]), d1 = c_muli<S>(c_sub(v[1], v[3]));
    v[0] = c_add(s0, s1);
    v[2] = c_sub(s0, s1);
    v[1] = c_add(d0, d1);
    v[3] = c_sub(d0, d1);
  }
};

template <int S> struct SmallDFT<8, S> {
  static __device__ __forceinline__ void run(float2* v) {
    const float kR = 0.70710678118654752440f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    SmallDFT<4, S>::run(e);
    SmallDFT<4, S>::run(o);
    // w8^1 = (1 + S i)/sqrt2, w8^2 = S i, w8^3 = (-1 + S i)/sqrt2
    // o1, o3 below are the products before the 1/sqrt2 scale
    const float2 o1 = c_add(o[1], c_muli<S>(o[1]));
    const float2 o2 = c_muli<S>(o[2]);
    const float2 o3 = c_sub(c_muli<S>(o[3]), o[3]);
    v[0] = c_add(e[0], o[0]);
    v[4] = c_sub(e[0], o[0]);
    v[1] = c_fma_s(o1, kR, e[1]);
    v[5] = c_fnma_s(o1, kR, e[1]);
    v[2] = c_add(e[2], o2);
    v[6] = c_sub(e[2], o2);
    v[3] = c_fma_s(o3, kR, e[3]);
    v[7] = c_fnma_s(o3, kR, e[3]);
  }
};

template <int S> struct SmallDFT<16, S> {
  static __device__ __forceinline__ void run(float2* v) {
    // cos / sin of k*pi/8, k = 0..7
    const float C[8] = {1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
                        0.0f, -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f};
    const float Sn[8] = {0.0f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f,
                         1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f};
    float2 e[8], o[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    SmallDFT<8, S>::run(e);
    SmallDFT<8, S>::run(o);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      float2 w = make_float2(C[k], S > 0 ? Sn[k] : -Sn[k]);
      float2 t = (k == 0) ? o[0] : c_mul(o[k], w);
      v[k] = c_add(e[k], t);
      v[k + 8] = c_sub(e[k], t);
    }
  }
};

// ---------------------------------------------------------------------------
// One Stockham pass of radix R over `nlines` lines.
//   in/out element (line, n) lives at [line*ls + n*es].
// LINE_FAST selects the thread->work mapping: consecutive threads take
// consecutive lines (best when ls == 1) or consecutive butterflies of one
// line (best when es == 1).  NLINES_CT > 0 makes the line count compile-time.
// ---------------------------------------------------------------------------
template <int N, int R, int NS, int S, bool LINE_FAST, int NLINES_CT>
__device__ __forceinline__ void stockham_pass(const float2* __restrict__ in, float2* __restrict__ out,
                                              int nlines_rt, int es, int ls, const float2* __restrict__ tw,
                                              int tid, int nthreads) {
  constexpr int NB = N / R;
  const int nlines = NLINES_CT > 0 ? NLINES_CT : nlines_rt;
  const int total = nlines * NB;
  for (int t = tid; t < total; t += nthreads) {
    int line, j;
    if constexpr (LINE_FAST) {
      if constexpr (NLINES_CT > 0) {
        line = t % NLINES_CT;
        j = t / NLINES_CT;
      } else {
        line = t % nlines;
        j = t / nlines;
      }
    } else {
      line = t / NB;
      j = t - line * NB;
    }
    const int k = j % NS;
    const float2* ip = in + line * ls;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = ip[(j + r * NB) * es];
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[k * r * (N / (NS * R))];
        if (S > 0) w.y = -w.y;
        v[r] = c_mul(v[r], w);
      }
    }
    SmallDFT<R, S>::run(v);
    const int base = (j / NS) * NS * R + k;
    float2* op = out + line * ls;
#pragma unroll
    for (int r = 0; r < R; ++r) op[(base + r * NS) * es] = v[r];
  }
}

template <int... Rs> struct RadixList {};

// Synchronisation between passes: the whole workgroup, or one wavefront
// when each wavefront owns its lines and its LDS slice.
struct BlockSync {
  static __device__ __forceinline__ void sync() { __syncthreads(); }
  static __device__ __forceinline__ int tid() { return threadIdx.x; }
  static __device__ __forceinline__ int count() { return blockDim.x; }
};
struct WaveSync {
  static __device__ __forceinline__ void sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  static __device__ __forceinline__ int tid() { return threadIdx.x & 63; }
  static __device__ __forceinline__ int count() { return 64; }
};

template <int N, int S, bool LINE_FAST, int NLINES_CT, int NS, class SY>
__device__ __forceinline__ float2* fft_passes(float2* a, float2*, int, int, int, const float2*, RadixList<>, SY) {
  return a;
}

// Runs all passes; returns the buffer holding the natural-order result.
// Caller must synchronise before (data in `a` visible); each pass ends with
// a sync so the result is visible on return.
template <int N, int S, bool LINE_FAST, int NLINES_CT, int NS, class SY, int R, int... Rest>
__device__ __forceinline__ float2* fft_passes(float2* a, float2* b, int nlines, int es, int ls, const float2* tw,
                                              RadixList<R, Rest...>, SY sy) {
  stockham_pass<N, R, NS, S, LINE_FAST, NLINES_CT>(a, b, nlines, es, ls, tw, SY::tid(), SY::count());
  SY::sync();
  return fft_passes<N, S, LINE_FAST, NLINES_CT, NS * R>(b, a, nlines, es, ls, tw, RadixList<Rest...>{}, sy);
}

template <int... Rs> struct RadixProduct;
template <> struct RadixProduct<> { static constexpr int value = 1; };
template <int R, int... Rest> struct RadixProduct<R, Rest...> {
  static constexpr int value = R * RadixProduct<Rest...>::value;
};

template <int... Rs> struct RadixCount { static constexpr int value = sizeof...(Rs); };

// A line-FFT plan: length N factored as Rs... (product must equal N).
template <int N_, int... Rs> struct Plan {
  static constexpr int N = N_;
  static constexpr int ZH = N_ / 2 + 1;
  static constexpr int NPASS = sizeof...(Rs);
  using radices = RadixList<Rs...>;
  static_assert(RadixProduct<Rs...>::value == N_, "radix product must equal N");
};

template <class P, int S, bool LINE_FAST, int NLINES_CT = 0>
__device__ __forceinline__ float2* lds_fft(float2* a, float2* b, int nlines, int es, int ls, const float2* tw) {
  return fft_passes<P::N, S, LINE_FAST, NLINES_CT, 1>(a, b, nlines, es, ls, tw, typename P::radices{}, BlockSync{});
}

// Same transform run by one wavefront on lines it owns.
template <class P, int S, bool LINE_FAST, int NLINES_CT = 0>
__device__ __forceinline__ float2* wave_fft(float2* a, float2* b, int nlines, int es, int ls, const float2* tw) {
  return fft_passes<P::N, S, LINE_FAST, NLINES_CT, 1>(a, b, nlines, es, ls, tw, typename P::radices{}, WaveSync{});
}

}  // namespace cdk
