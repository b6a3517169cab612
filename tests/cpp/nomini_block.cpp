// The TranRotaConfSearchNoMini.cu:156-487 block as a C++ caller writes it
// against include/cdk.h and include/cdk_reference.hpp (the binding shown in
// INTEGRATION.md), built by tests/cpp/Makefile and run by
// tests/test_cpp_boundary.py.
//
//   nomini_block host <in> <out>
//       reference-signature helpers that run on the host (Rotate,
//       GetMinCoors, GetMaxCoors, GetIdxOfAtomsForVdwRadius,
//       QuaternionUniformSampling) on the inputs in <in>, results to <out>;
//       then the status-code paths of the C-ABI (bad counts, bad arguments,
//       a context on a device that does not exist).
//   nomini_block search <out>
//       the driver's search loop on the GPU: synthetic 64^3 protein, a
//       40-atom ligand, 3 conformers; per conformer QuaternionUniformSampling
//       (caller's std::mt19937_64), FilterQuaternions, cdk_search,
//       cdk_rank_poses_quats (top 5) and FillLigandGrid of the best pose;
//       results to <out>.
//
// Files are flat little-endian records: int32 count, then that many values.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "cdk.h"
#include "cdk_reference.hpp"

#define CDK_CALL(x)                                                         \
  do {                                                                      \
    if ((x) != CDK_OK) {                                                    \
      std::fprintf(stderr, "%s failed: %s\n", #x, cdk_last_error());       \
      return 1;                                                             \
    }                                                                       \
  } while (0)

namespace {

struct Reader {
  FILE* f;
  template <class T>
  std::vector<T> next() {
    int n = 0;
    if (std::fread(&n, 4, 1, f) != 1 || n < 0) return {};
    std::vector<T> v(n);
    if (n && std::fread(v.data(), sizeof(T), n, f) != (size_t)n) v.clear();
    return v;
  }
};

struct Writer {
  FILE* f;
  template <class T>
  void put(const T* p, size_t n) {
    const int c = (int)n;
    std::fwrite(&c, 4, 1, f);
    if (n) std::fwrite(p, sizeof(T), n, f);
  }
  template <class T>
  void put(const std::vector<T>& v) {
    put(v.data(), v.size());
  }
};

int host_mode(const char* in_path, const char* out_path) {
  FILE* fi = std::fopen(in_path, "rb");
  FILE* fo = std::fopen(out_path, "wb");
  if (!fi || !fo) return 2;
  Reader r{fi};
  Writer w{fo};
  std::vector<float> quats = r.next<float>(), coords = r.next<float>(), atom_radii = r.next<float>(),
                     probe = r.next<float>();
  std::vector<long long> seed_n = r.next<long long>();  // seed, calls, quaternions per call
  std::fclose(fi);
  const int nq = (int)quats.size() / 4, na = (int)coords.size() / 3;
  // Rotate every (quaternion, atom) as NoMini:352-358 does
  std::vector<float> rot((size_t)nq * na * 3);
  for (int i = 0; i < nq; ++i)
    for (int a = 0; a < na; ++a) Rotate(&quats[i * 4], &coords[a * 3], &rot[((size_t)i * na + a) * 3]);
  std::vector<float> mn(nq * 3), mx(nq * 3);
  GetMinCoors(nq, na, rot.data(), mn.data());
  GetMaxCoors(nq, na, rot.data(), mx.data());
  w.put(rot);
  w.put(mn);
  w.put(mx);
  // channel assignment
  const int nv = (int)probe.size(), nr = (int)atom_radii.size();
  int used_n = -1;
  std::vector<int> used;
  std::vector<std::vector<int> > lists(nv);
  GetIdxOfAtomsForVdwRadius(nr, atom_radii.data(), nv, probe.data(), used_n, used, lists);
  std::vector<int> ari(nr, -1);
  for (int v = 0; v < nv; ++v)
    for (int a : lists[v]) ari[a] = v;
  w.put(&used_n, 1);
  w.put(used);
  w.put(ari);
  // successive QuaternionUniformSampling calls on one generator
  std::mt19937_64 gen((unsigned long long)seed_n[0]);
  float* q = nullptr;
  std::vector<float> all;
  for (long long c = 0; c < seed_n[1]; ++c) {
    QuaternionUniformSampling(gen, q, (int)seed_n[2]);
    all.insert(all.end(), q, q + 4 * seed_n[2]);
  }
  delete[] q;
  w.put(all);
  // status codes instead of exits
  std::vector<int> codes;
  codes.push_back(cdk_rotate_host(-1, 3, nullptr, nullptr, nullptr));
  codes.push_back(cdk_min_max_coors(1, 0, rot.data(), nullptr, nullptr));
  cdk_context* bad = nullptr;
  codes.push_back(cdk_context_create(1 << 20, &bad));
  codes.push_back(bad == nullptr ? 0 : 1);
  codes.push_back(cdk_idx_of_atoms_for_vdw_radius(-1, nullptr, 0, nullptr, nullptr, nullptr, nullptr));
  codes.push_back((int)std::strlen(cdk_last_error()) > 0);
  w.put(codes);
  std::fclose(fo);
  return 0;
}

int search_mode(const char* out_path) {
  FILE* fo = std::fopen(out_path, "wb");
  if (!fo) return 2;
  Writer w{fo};
  // inputs: the single-ligand workload's protein and ligand recipes
  const int N = 64, nch = 9, natom = 40, nconf = 3, numOfRotaSample = 300, nLowest = 5;
  const float spacing = 0.5f;
  std::vector<float> values((size_t)nch * N * N * N), radii(8), mid(3), len(3);
  CDK_CALL(cdk_synth_protein(12, 3000, 32.0f, N, spacing, values.data(), radii.data(), mid.data(), len.data()));
  std::vector<float> x(natom * 3), charge(natom), eps(natom), rad(natom), confs((size_t)nconf * natom * 3);
  CDK_CALL(cdk_synth_ligand(102, natom, x.data(), charge.data(), eps.data(), rad.data()));
  CDK_CALL(cdk_synth_conformers(102 + 7919, natom, x.data(), nconf, confs.data()));

  // NoMini:156-252 set-up
  cdk_context* ctx;
  CDK_CALL(cdk_context_create(0, &ctx));
  cdk_protein* prot;
  CDK_CALL(cdk_protein_create(ctx, nch, N, N, N, mid.data(), len.data(), spacing, radii.data(), values.data(), &prot));
  cdk_ligand* lig;
  CDK_CALL(cdk_ligand_create(prot, natom, charge.data(), eps.data(), rad.data(), &lig));
  std::mt19937_64 gen(202);
  float* quaters = nullptr;
  float* qUsed = nullptr;
  for (int c = 0; c < nconf; ++c) {
    float* conformerCoor = &confs[(size_t)c * natom * 3];
    QuaternionUniformSampling(gen, quaters, numOfRotaSample);
    const int nUsed = FilterQuaternions(conformerCoor, natom, numOfRotaSample, quaters, len[0], len[1], len[2],
                                        numOfRotaSample, qUsed);
    std::vector<float> minE(nUsed), minc(3 * (size_t)nUsed);
    std::vector<int> idx(3 * (size_t)nUsed);
    CDK_CALL(cdk_search(ctx, prot, lig, 1, conformerCoor, nUsed, qUsed, 0, CDK_ENGINE_FUSED, minE.data(), idx.data(),
                        minc.data()));
    // NoMini:466-487: top poses from the conformer and the rotations
    std::vector<int> rank(nLowest);
    std::vector<double> pose((size_t)nLowest * natom * 3);
    CDK_CALL(cdk_rank_poses_quats(prot, natom, nUsed, conformerCoor, qUsed, minE.data(), idx.data(), minc.data(),
                                  nLowest, rank.data(), pose.data()));
    // the best pose's ligand grids through the reference signature
    std::vector<float> rot((size_t)natom * 3);
    for (int a = 0; a < natom; ++a) Rotate(&qUsed[rank[0] * 4], &conformerCoor[a * 3], &rot[a * 3]);
    float mn[3], mx[3];
    GetMinCoors(1, natom, rot.data(), mn);
    GetMaxCoors(1, natom, rot.data(), mx);
    int nu = 0;
    std::vector<int> usedIdx;
    std::vector<std::vector<int> > lists(8);
    GetIdxOfAtomsForVdwRadius(natom, rad.data(), 8, radii.data(), nu, usedIdx, lists);
    int box = 2;  // every atom's cell and its +1 corner inside (cpuNaive.cpp:103-105's int(len/h) + 2, plus one)
    for (int k = 0; k < 3; ++k) box = std::max(box, (int)((mx[k] - mn[k]) / spacing) + 3);
    std::vector<float> grid((size_t)(nu + 1) * box * box * box, 0.0f);
    FillLigandGrid(1, natom, rot.data(), mn, charge.data(), eps.data(), nu, usedIdx, lists, box, box, box, spacing,
                   grid.data());
    double mass = 0;
    for (float v : grid) mass += v;
    w.put(std::vector<float>(qUsed, qUsed + 4 * (size_t)nUsed));
    w.put(minE);
    w.put(idx);
    w.put(minc);
    w.put(rank);
    w.put(pose);
    w.put(&mass, 1);
  }
  delete[] quaters;
  delete[] qUsed;
  CDK_CALL(cdk_ligand_destroy(lig));
  CDK_CALL(cdk_protein_destroy(prot));
  CDK_CALL(cdk_context_destroy(ctx));
  std::fclose(fo);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 4 && std::string(argv[1]) == "host") return host_mode(argv[2], argv[3]);
  if (argc == 3 && std::string(argv[1]) == "search") return search_mode(argv[2]);
  std::fprintf(stderr, "usage: nomini_block host <in> <out> | search <out>\n");
  return 2;
}
