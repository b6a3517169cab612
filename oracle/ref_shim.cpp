// ORACLE - test infrastructure only.
//
// extern "C" entry points over the reference's own pure-math sources
// (Rotate.cpp, QuaternionMultiply.cpp, GetMinCoors.cpp, GetMaxCoors.cpp,
// GetIdxOfAtomsForVdwRadius.cpp, FillLigandGrid.cpp,
// QuaternionUniformSampling.cpp, FilterQuaternions.cpp, ReadGrids.cpp),
// compiled where they lie by oracle/Makefile into oracle/_ref/.  Used only to
// write and check golden vectors; this file contains no reference code.
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "FillLigandGrid.h"
#include "FilterQuaternions.h"
#include "GetIdxOfAtomsForVdwRadius.h"
#include "GetMaxCoors.h"
#include "GetMinCoors.h"
#include "QuaternionUniformSampling.h"
#include "ReadGrids.h"
#include "Rotate.h"

extern "C" {

void ref_rotate_batch(int nq, int natom, const float* quats, const float* coords, float* out) {
  for (int i = 0; i < nq; ++i)
    for (int a = 0; a < natom; ++a) Rotate(quats + i * 4, coords + a * 3, out + ((size_t)i * natom + a) * 3);
}

void ref_min_coors(int nq, int natom, float* coors, float* out) { GetMinCoors(nq, natom, coors, out); }
void ref_max_coors(int nq, int natom, float* coors, float* out) { GetMaxCoors(nq, natom, coors, out); }

void ref_idx_of_atoms_for_vdw_radius(int natom, float* radii, int nv, float* grid_radii, int* num_used,
                                     int* idx_used, int* atom_radius_index) {
  int nu = 0;
  std::vector<int> used;
  std::vector<std::vector<int> > lists(nv);
  GetIdxOfAtomsForVdwRadius(natom, radii, nv, grid_radii, nu, used, lists);
  *num_used = nu;
  for (int i = 0; i < nu; ++i) idx_used[i] = used[i];
  for (int r = 0; r < nv; ++r)
    for (size_t k = 0; k < lists[r].size(); ++k) atom_radius_index[lists[r][k]] = r;
}

void ref_fill_ligand_grid(int nq, int natom, float* coors, float* mincoors, float* charge, float* eps,
                          int num_used, const int* idx_used, int nv, const int* radius_offsets,
                          const int* radius_atoms, int X, int Y, int Z, float spacing, float* grid) {
  std::vector<int> used(idx_used, idx_used + num_used);
  std::vector<std::vector<int> > lists(nv);
  for (int r = 0; r < nv; ++r)
    for (int i = radius_offsets[r]; i < radius_offsets[r + 1]; ++i) lists[r].push_back(radius_atoms[i]);
  FillLigandGrid(nq, natom, coors, mincoors, charge, eps, num_used, used, lists, X, Y, Z, spacing, grid);
}

void ref_quaternion_uniform_sampling(uint64_t seed, int ncall, int n, float* out) {
  std::mt19937_64 gen(seed);
  float* q = 0;
  for (int c = 0; c < ncall; ++c) {
    QuaternionUniformSampling(gen, q, n);
    std::memcpy(out + (size_t)c * n * 4, q, sizeof(float) * 4 * n);
  }
  delete[] q;
}

int ref_filter_quaternions(float* coor, int natom, int nq, float* quats, float lx, float ly, float lz,
                           int max_used, float* out) {
  float* used = 0;
  const int n = FilterQuaternions(coor, natom, nq, quats, lx, ly, lz, max_used, used);
  std::memcpy(out, used, sizeof(float) * 4 * n);
  delete[] used;
  return n;
}

// ReadGrids.cpp:6-68 into caller buffers: header fields, up to max_radii
// radii and up to max_values values (returns -1 if either would overflow).
int ref_read_grids(const char* path, int* num_grids, int* dims, float* mid, float* len, float* spacing,
                   float* restraint_k, float* radii, int max_radii, float* values, long long max_values) {
  int ng = 0, x = 0, y = 0, z = 0;
  float mx = 0, my = 0, mz = 0, lx = 0, ly = 0, lz = 0, h = 0, rk = 0;
  float* r = 0;
  float* v = 0;
  ReadGrids(ng, x, y, z, mx, my, mz, lx, ly, lz, h, rk, r, v, std::string(path));
  *num_grids = ng;
  dims[0] = x;
  dims[1] = y;
  dims[2] = z;
  mid[0] = mx;
  mid[1] = my;
  mid[2] = mz;
  len[0] = lx;
  len[1] = ly;
  len[2] = lz;
  *spacing = h;
  *restraint_k = rk;
  const long long nv = (long long)x * y * z * ng;
  int rc = 0;
  if (ng - 1 > max_radii || nv > max_values) {
    rc = -1;
  } else {
    if (ng > 1) std::memcpy(radii, r, sizeof(float) * (ng - 1));
    if (nv > 0) std::memcpy(values, v, sizeof(float) * nv);
  }
  delete[] r;
  delete[] v;
  return rc;
}

}  // extern "C"
