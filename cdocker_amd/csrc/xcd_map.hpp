// Workgroup-to-XCD maps of the kz-plane and line-group kernels.  MI355X
// dispatches the workgroups of a launch to its 8 XCDs round-robin
// (workgroup w to XCD w % 8), and each XCD has its own L2; these maps pick
// which (item, unit) a workgroup computes so each XCD works on its own share
// of the items (and of their protein data) and every XCD gets the same work.
// Host and device: the kernels call them, and a host test checks that every
// (item, unit) pair is covered exactly once.
#pragma once
#include <hip/hip_runtime.h>

namespace cdk {

// Workgroup w of a launch over G units (pose groups, or (pose group, block
// row) pairs) times M = 8 Q + R items (kz planes, or line groups of the
// spectra).  Workgroups go to the 8 XCDs round-robin (w % 8), so XCD x runs
// items [Q x, Q x + Q) of every unit -- its L2 holds its own eighth of the
// protein data -- and, of each block of 8 units, R of the 8 R leftover items
// (item 8 Q + i of unit g8 goes to XCD (g8 R + i) % 8): every XCD gets the
// same work.  (Padding the items to a multiple of 8 instead left some XCDs a
// whole item per unit more than others, and those set the kernel's time.)
// Launch xcd_unit_grid(Q, R, G) workgroups; false for those past the last
// unit.
__host__ __device__ __forceinline__ unsigned xcd_unit_grid(int Q, int R, int G) {
  return 8u * (unsigned)(8 * Q + R) * (unsigned)((G + 7) / 8);
}

__host__ __device__ __forceinline__ bool xcd_unit_of(int w, int Q, int R, int G, int& item, int& unit) {
  const int x = w & 7, m = w >> 3;
  const int per = 8 * Q + R;
  const int b = m / per, r = m - b * per;
  int g8;
  if (r < 8 * Q) {
    g8 = r / Q;
    item = Q * x + (r - g8 * Q);
  } else {
    const int t = x + 8 * (r - 8 * Q);
    g8 = t / R;
    item = 8 * Q + (t - g8 * R);
  }
  unit = 8 * b + g8;
  return unit < G;
}

// Plane-major form for kernels whose items share large per-item data (the
// protein spectra of a kz plane): XCD x runs its planes Qx ... Qx + Q - 1
// one after another, all G units of a plane before the next, then its share
// of the R leftover planes' items (spread over the XCDs in turn).  The
// workgroups resident on an XCD at a time then cover one or two planes, whose
// spectra its L2 holds.
__host__ __device__ __forceinline__ unsigned xcd_plane_grid(int Q, int R, int G) {
  return 8u * (unsigned)(Q * G + (R * G + 7) / 8);
}
__host__ __device__ __forceinline__ bool xcd_plane_of(int w, int Q, int R, int G, int& item, int& unit) {
  const int x = w & 7, m = w >> 3;
  if (m < Q * G) {
    const int p = m / G;
    item = Q * x + p;
    unit = m - p * G;
    return true;
  }
  const int j = (m - Q * G) * 8 + x;
  if (j >= R * G) return false;
  const int i = j / G;
  item = 8 * Q + i;
  unit = j - i * G;
  return true;
}

}  // namespace cdk
