// Entry points that allocate nothing: device buffers in, device buffers out,
// working memory in a fixed per-device array (noalloc.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace cdk {

// Poses one call may carry (the fixed array holds their window metadata and
// argmin keys: 72 B each).
constexpr int kNoAllocMaxPoses = 16384;

// Looks up the device address of this device's fixed array (the device must
// be current); done once per device.
hipError_t noalloc_init(int device);

// NoMini:441-464 windowed argmin of nq energy grids [nq][n][n][n] (device),
// windows [nq][3] (device): energy [nq] and index [nq][3] (device), enqueued
// on st; nothing is allocated and nothing waits.  The device must be
// current.  Calls on one device share the fixed array: calls that overlap in
// time on one device race.
hipError_t argmin_window_noalloc(int device, hipStream_t st, int nq, int n, const float* grids, const int* windows,
                                 float* energy, int* index);

}  // namespace cdk
