/* cdk_reference.hpp - the reference's host helper functions, with their
 * exact C++ signatures, over libcdk.so (include/cdk.h).
 *
 * A driver written against the reference's headers (Rotate.h,
 * GetMinCoors.h, GetMaxCoors.h, GetIdxOfAtomsForVdwRadius.h,
 * QuaternionUniformSampling.h, FilterQuaternions.h, FillLigandGrid.h)
 * includes this header instead and links -lcdk; the calls are unchanged.
 * Header-only, C++11.
 *
 * Behaviour kept: outputs passed as float*& are freed and reallocated with
 * new[] (QuaternionUniformSampling.cpp:8-9, FilterQuaternions.cpp:67-68);
 * GetIdxOfAtomsForVdwRadius appends to the caller's vectors
 * (GetIdxOfAtomsForVdwRadius.cpp:21, :31); FillLigandGrid adds into the
 * caller's zeroed grid (FillLigandGrid.cpp:66).  On an error the reference's
 * helpers print and return; these print cdk_last_error() to std::cout and
 * return with the outputs untouched.
 *
 * FillLigandGrid and FilterQuaternions run on the GPU through one context
 * per process (device $CDK_DEVICE, default 0), created on first use.
 * FilterQuaternions' over-cap selection is a shuffle seeded from std::rand()
 * (the reference shuffles with std::random_shuffle, which draws from rand()).
 */
#ifndef CDK_REFERENCE_HPP
#define CDK_REFERENCE_HPP

#include <cstdlib>
#include <iostream>
#include <random>
#include <vector>

#include "cdk.h"

namespace cdk_reference_detail {

inline bool ok(int rc, const char* what) {
  if (rc == CDK_OK) return true;
  std::cout << what << ": " << cdk_last_error() << std::endl;
  return false;
}

// The process-wide context of the GPU-backed helpers.
inline cdk_context* context() {
  static cdk_context* ctx = [] {
    const char* d = std::getenv("CDK_DEVICE");
    cdk_context* c = nullptr;
    if (!ok(cdk_context_create(d ? std::atoi(d) : 0, &c), "cdk_context_create")) return (cdk_context*)nullptr;
    return c;
  }();
  return ctx;
}

}  // namespace cdk_reference_detail

/* Rotate.h:7-8 (Rotate.cpp:4-27). */
inline void Rotate(const float* const quaternion, const float* const icoor, float* const fcoor) {
  cdk_reference_detail::ok(cdk_rotate_host(1, 1, quaternion, icoor, fcoor), "Rotate");
}

/* GetMinCoors.h:7 (GetMinCoors.cpp:4-29). */
inline void GetMinCoors(int nQuaternions, int nAtoms, float* coors, float* mincoors) {
  cdk_reference_detail::ok(cdk_min_max_coors(nQuaternions, nAtoms, coors, mincoors, nullptr), "GetMinCoors");
}

/* GetMaxCoors.h:7 (GetMaxCoors.cpp:3-27). */
inline void GetMaxCoors(int nQuaternions, int nAtoms, float* coors, float* maxcoors) {
  cdk_reference_detail::ok(cdk_min_max_coors(nQuaternions, nAtoms, coors, nullptr, maxcoors), "GetMaxCoors");
}

/* GetIdxOfAtomsForVdwRadius.h:8-11 (GetIdxOfAtomsForVdwRadius.cpp:4-35):
 * atom i goes to idxOfAtomVdwRadius[nearest radius] (the caller sizes it
 * numOfVdwGrids); the used radii are appended to idxOfVdwUsed. */
inline void GetIdxOfAtomsForVdwRadius(int nAtoms, float* atomRadii, int numOfVdwGrids, float* gridRadii,
                                      int& numOfVdwGridsUsed, std::vector<int>& idxOfVdwUsed,
                                      std::vector<std::vector<int> >& idxOfAtomVdwRadius) {
  std::vector<int> used(numOfVdwGrids > 0 ? numOfVdwGrids : 1), ari(nAtoms > 0 ? nAtoms : 1);
  int nu = 0;
  if (!cdk_reference_detail::ok(cdk_idx_of_atoms_for_vdw_radius(nAtoms, atomRadii, numOfVdwGrids, gridRadii, &nu,
                                                                 used.data(), ari.data()),
                                "GetIdxOfAtomsForVdwRadius"))
    return;
  for (int i = 0; i < nAtoms; ++i) idxOfAtomVdwRadius[ari[i]].push_back(i);
  numOfVdwGridsUsed = 0;
  for (int r = 0; r < numOfVdwGrids; ++r)
    if (!idxOfAtomVdwRadius[r].empty()) {
      numOfVdwGridsUsed += 1;
      idxOfVdwUsed.push_back(r);
    }
}

/* QuaternionUniformSampling.h:10 (QuaternionUniformSampling.cpp:5-24): the
 * same draws from the caller's generator. */
inline void QuaternionUniformSampling(std::mt19937_64& gen, float*& quaters, int nquaters) {
  delete[] quaters;
  quaters = new float[4 * (nquaters > 0 ? nquaters : 0)];
  cdk_reference_detail::ok(cdk_quaternion_uniform_sampling_mt(&gen, nquaters, quaters), "QuaternionUniformSampling");
}

/* FilterQuaternions.h:7-10 (FilterQuaternions.cpp:11-83): the rotations
 * whose extent is below the grid length on every axis, at most
 * maxNQuaternionsUsed of them; returns their count. */
inline int FilterQuaternions(float* coor, int nAtom, int nQuater, float* quaternions, float gridLenX, float gridLenY,
                             float gridLenZ, int maxNQuaternionsUsed, float*& quaternionsUsed) {
  cdk_context* ctx = cdk_reference_detail::context();
  if (!ctx) return 0;
  std::vector<float> kept(4 * (size_t)(nQuater > 0 ? nQuater : 1));
  int n = 0;
  const uint64_t seed = (uint64_t)std::rand();
  if (!cdk_reference_detail::ok(cdk_filter_quaternions(ctx, coor, nAtom, nQuater, quaternions, gridLenX, gridLenY,
                                                       gridLenZ, maxNQuaternionsUsed, seed, &n, kept.data()),
                                "FilterQuaternions"))
    return 0;
  delete[] quaternionsUsed;
  quaternionsUsed = new float[4 * (size_t)n];
  for (int i = 0; i < 4 * n; ++i) quaternionsUsed[i] = kept[i];
  return n;
}

/* FillLigandGrid.h:7-14 (FillLigandGrid.cpp:7-120): adds the ligand grids
 * [nQuaternions][numOfVdwGridsUsed + 1][xdim][ydim][zdim] into
 * ligandGridValues (the caller zeroes it first, as NoMini:389 does). */
inline void FillLigandGrid(int nQuaternions, int nAtoms, float* coors, float* mincoors, float* charge, float* epsilon,
                           int numOfVdwGridsUsed, std::vector<int> idxOfVdwUsed,
                           std::vector<std::vector<int> >& idxOfAtomVdwRadius, int xdim, int ydim, int zdim,
                           float spacing, float* ligandGridValues) {
  cdk_context* ctx = cdk_reference_detail::context();
  if (!ctx) return;
  const int nv = (int)idxOfAtomVdwRadius.size();
  std::vector<int> offs(nv + 1, 0), atoms;
  for (int r = 0; r < nv; ++r) {
    offs[r] = (int)atoms.size();
    atoms.insert(atoms.end(), idxOfAtomVdwRadius[r].begin(), idxOfAtomVdwRadius[r].end());
  }
  offs[nv] = (int)atoms.size();
  if (atoms.empty()) atoms.push_back(0);
  if (idxOfVdwUsed.empty()) idxOfVdwUsed.push_back(0);
  cdk_reference_detail::ok(cdk_fill_ligand_grid(ctx, nQuaternions, nAtoms, coors, mincoors, charge, epsilon,
                                                numOfVdwGridsUsed, idxOfVdwUsed.data(), nv, offs.data(), atoms.data(),
                                                xdim, ydim, zdim, spacing, ligandGridValues),
                           "FillLigandGrid");
}

#endif /* CDK_REFERENCE_HPP */
