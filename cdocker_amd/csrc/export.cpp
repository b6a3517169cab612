// Retained-pose export and the grid handoff to the host minimiser.
//
// Pose files and energy rows follow the reference drivers' output: one PDB
// per retained pose named conformer_<c>_<rank>.pdb and one row
// "file,conformer,rank,energy" per pose in energy.txt
// (TranRotaConfSearchNoMini.cu:489-502).  The reference writes the PDB with
// OpenBabel from the full molecule; here the caller passes atom names and
// elements and the file holds HETATM records only.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "engine.hpp"
#include "rescore.hpp"

namespace cdk {

namespace {

std::string pdb_name(const char* name) {
  // Atom names shorter than 4 characters start in column 14 (PDB convention).
  std::string n = name ? name : "C";
  if (n.size() > 4) n.resize(4);
  if (n.size() < 4) n = " " + n;
  while (n.size() < 4) n += ' ';
  return n;
}

}  // namespace

std::string pose_file_name(int conformer, int rank) {
  return "conformer_" + std::to_string(conformer) + "_" + std::to_string(rank) + ".pdb";
}

std::string energy_row(int conformer, int rank, float energy) {
  // std::ostream's default float formatting, as the drivers' energyFile <<.
  std::ostringstream os;
  os << pose_file_name(conformer, rank) << "," << conformer << "," << rank << "," << energy;
  return os.str();
}

void write_pose_pdb(const std::string& path, int natom, const double* xyz, const char* const* names,
                    const char* const* elements) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) throw Error(kErrInvalid, "cannot open " + path + " for writing");
  for (int a = 0; a < natom; ++a) {
    const std::string nm = pdb_name(names ? names[a] : nullptr);
    const char* el = elements ? elements[a] : "C";
    std::fprintf(f, "HETATM%5d %4s LIG A   1    %8.3f%8.3f%8.3f  1.00  0.00          %2.2s\n", a + 1, nm.c_str(),
                 xyz[a * 3 + 0], xyz[a * 3 + 1], xyz[a * 3 + 2], el);
  }
  std::fprintf(f, "END\n");
  if (std::fclose(f) != 0) throw Error(kErrInvalid, "write failed for " + path);
}

void write_poses(const char* dir, int conformer, int n, int natom, const double* poses, const float* energy,
                 const char* const* names, const char* const* elements, const char* energy_path) {
  const std::string d = (dir && *dir) ? std::string(dir) + "/" : std::string();
  std::string rows;
  for (int i = 0; i < n; ++i) {
    write_pose_pdb(d + pose_file_name(conformer, i), natom, poses + (size_t)i * natom * 3, names, elements);
    rows += energy_row(conformer, i, energy[i]) + "\n";
  }
  if (energy_path && *energy_path) {
    FILE* f = std::fopen(energy_path, "a");
    if (!f) throw Error(kErrInvalid, std::string("cannot open ") + energy_path + " for appending");
    std::fputs(rows.c_str(), f);
    if (std::fclose(f) != 0) throw Error(kErrInvalid, std::string("write failed for ") + energy_path);
  }
}

void grid_for_openmm(Context& ctx, const Protein& prot, int channel, double* values, double* bounds_nm) {
  std::lock_guard<std::mutex> lock(ctx.mu);
  hip_check(hipSetDevice(ctx.device), "hipSetDevice");
  const int N = prot.N;
  const size_t G = prot.G();
  if (values) {
    DevBuf out;
    out.ensure(G * sizeof(double));
    hip_check(launch_grid_xfast(ctx.stream, prot.grid.as<float>() + (size_t)channel * G, N, N, N, out.as<double>()),
              "grid order");
    hip_check(hipMemcpyAsync(values, out.p, G * sizeof(double), hipMemcpyDeviceToHost, ctx.stream), "grid D2H");
    hip_check(hipStreamSynchronize(ctx.stream), "grid sync");
  }
  if (bounds_nm) {
    // TranRotaConfSearchGrid.cu:149-154, scaled by NmPerAngstrom as
    // AddGridForcesToOpenMMSystem.cpp:45-50 passes them.
    for (int k = 0; k < 3; ++k) {
      const double gmin = prot.grid_min(k);
      const double gmax = gmin + (double)((float)(N - 1) * prot.h);
      bounds_nm[k] = gmin * 0.1;
      bounds_nm[3 + k] = gmax * 0.1;
    }
  }
}

}  // namespace cdk
