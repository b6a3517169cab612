// Host-side input producers (see host_inputs.cpp).
#pragma once
#include <cstdint>
#include <random>
#include <string>
#include <vector>

namespace cdk {

struct Xoshiro {
  uint64_t s[4];
  explicit Xoshiro(uint64_t seed);
  uint64_t next();
  double uniform();
  double normal();
};

extern const float kProbeRadii[8];

void quaternion_uniform_sampling(std::mt19937_64& gen, int n, float* out);
// Rotate.cpp:4-27 on the host, the same FP32 operations in the same order as
// the device's (contraction off): out [nq][natom][3].
void rotate_host(int nq, int natom, const float* quats, const float* coords, float* out);
// GetMinCoors.cpp:4-29 / GetMaxCoors.cpp:3-27 over rotated coordinates
// [nq][natom][3]: the first atom's value, replaced only by a strictly
// smaller (larger) one.  Either output may be null.
void min_max_coors(int nq, int natom, const float* coors, float* mincoors, float* maxcoors);
// The atoms behind synth_protein: positions uniform in [0, box)^3, charge
// N(0, 0.3), radius from the probe set, epsilon uniform [0.05, 0.25].
void synth_protein_atoms(uint64_t seed, int natom, float box, double* xyz, double* charge, double* epsilon,
                         double* radius);
void synth_protein(uint64_t seed, int natom, float box, int n, float h, float* values, float* radii, float* mid3,
                   float* len3);
void synth_ligand(uint64_t seed, int natom, float* coords, float* charge, float* eps, float* radius);
void synth_conformers(uint64_t seed, int natom, const float* coords, int nconf, float* out);

struct GridHeader {
  int num_grids = 0;
  int dims[3] = {0, 0, 0};
  float mid[3] = {0, 0, 0};
  float len[3] = {0, 0, 0};
  float spacing = 0, restraint_k = 0;
  std::vector<float> radii;
};

void read_grid_header(const std::string& path, GridHeader& g);
void read_grids(const std::string& path, GridHeader& g, std::vector<float>& values);
void write_grids(const std::string& path, const GridHeader& g, const float* values);

}  // namespace cdk
