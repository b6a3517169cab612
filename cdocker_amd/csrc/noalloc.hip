// The windowed argmin over energy grids (NoMini:441-464) for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops): the per-pose window metadata and argmin keys live in a fixed
// file-scope device array instead of library allocations.
#include "noalloc.hpp"

#include <hip/hip_runtime.h>

#include <mutex>

#include "cdk_types.hpp"
#include "kernels.hpp"

namespace cdk {

namespace {

constexpr size_t kMetaBytes = (size_t)kNoAllocMaxPoses * sizeof(RotMeta);
constexpr size_t kScratchBytes = kMetaBytes + (size_t)kNoAllocMaxPoses * sizeof(unsigned long long);

}  // namespace

// [RotMeta x kNoAllocMaxPoses][key x kNoAllocMaxPoses], 1.2 MB; one copy per
// device, held until the process ends.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// The device address of this device's copy of `scratch` (the device must be
// current).  The array's name in host code is only a host-side handle.
void* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch[device] = p;
  }
  return g_scratch[device];
}

// Window metadata from the caller's windows (clamped to the grid edge, valid
// when non-empty) and keys set to the empty key.
__global__ void k_noalloc_prep(const int* __restrict__ windows, int nq, int n, RotMeta* __restrict__ meta,
                               unsigned long long* __restrict__ keys) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  RotMeta m;
  int ok = 1;
  for (int k = 0; k < 3; ++k) {
    m.minc[k] = 0.f;
    m.len[k] = 0.f;
    m.box[k] = 0;
    m.win[k] = min(windows[q * 3 + k], n);
    ok = ok && m.win[k] > 0;
  }
  m.valid = ok;
  m.pad = 0;
  meta[q] = m;
  keys[q] = ~0ull;
}

// Keys to (energy, index): the inverse of the argmin key encoding (FP32
// energy bits made order-preserving in the high word, linear index low).
__global__ void k_noalloc_decode(const unsigned long long* __restrict__ keys, int nq, int n,
                                 float* __restrict__ energy, int* __restrict__ index) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const unsigned long long key = keys[q];
  if (key == ~0ull) {
    energy[q] = __int_as_float(0x7f800000);
    index[q * 3] = index[q * 3 + 1] = index[q * 3 + 2] = -1;
    return;
  }
  unsigned u = (unsigned)(key >> 32);
  u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  energy[q] = __uint_as_float(u);
  const unsigned lin = (unsigned)(key & 0xffffffffu);
  index[q * 3 + 0] = (int)(lin / (unsigned)(n * n));
  index[q * 3 + 1] = (int)((lin / (unsigned)n) % (unsigned)n);
  index[q * 3 + 2] = (int)(lin % (unsigned)n);
}

}  // namespace

hipError_t noalloc_init(int device) { return scratch_ptr(device) ? hipSuccess : hipErrorInvalidSymbol; }

hipError_t argmin_window_noalloc(int device, hipStream_t st, int nq, int n, const float* grids, const int* windows,
                                 float* energy, int* index) {
  if (nq < 0 || nq > kNoAllocMaxPoses || n <= 0) return hipErrorInvalidValue;
  if (nq == 0) return hipSuccess;
  unsigned char* base = static_cast<unsigned char*>(scratch_ptr(device));
  if (!base) return hipErrorInvalidSymbol;
  RotMeta* meta = reinterpret_cast<RotMeta*>(base);
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(base + kMetaBytes);
  const unsigned blocks = (unsigned)((nq + 255) / 256);
  hipLaunchKernelGGL(k_noalloc_prep, dim3(blocks), dim3(256), 0, st, windows, nq, n, meta, keys);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_argmin_full(st, grids, meta, keys, n, nq);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_noalloc_decode, dim3(blocks), dim3(256), 0, st, keys, nq, n, energy, index);
  return hipGetLastError();
}

}  // namespace cdk
