// extern "C" boundary (include/cdk.h).  Every entry point converts C++
// exceptions into status codes and a thread-local message.
#include "../../include/cdk.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <random>
#include <string>

#include "engine.hpp"
#include "host_inputs.hpp"
#include "kernels.hpp"
#include "noalloc.hpp"
#include "server.hpp"

struct cdk_context {
  std::unique_ptr<cdk::Context> c;
};
struct cdk_protein {
  cdk::Protein p;
};
struct cdk_ligand {
  std::unique_ptr<cdk::Ligand> l;
  const cdk_protein* prot = nullptr;
};
struct cdk_batch {
  cdk::Batch b;
  const cdk_protein* prot = nullptr;
};
struct cdk_server {
  std::unique_ptr<cdk::Server> s;
};
struct cdk_rng {
  std::mt19937_64 g;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <class F>
int guard(F&& f) {
  try {
    f();
    return CDK_OK;
  } catch (const cdk::Error& e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(CDK_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(CDK_ERR_INVALID, e.what());
  }
}

void require(bool ok, const char* msg) {
  if (!ok) throw cdk::Error(CDK_ERR_INVALID, msg);
}

template <class T>
struct Tmp {
  cdk::DevBuf b;
  explicit Tmp(size_t n) { b.ensure(std::max<size_t>(n, 1) * sizeof(T)); }
  T* p() { return b.as<T>(); }
};

void h2d(cdk::Context& c, void* d, const void* h, size_t bytes) {
  if (bytes) cdk::hip_check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c.stream), "H2D");
}
void d2h(cdk::Context& c, void* h, const void* d, size_t bytes) {
  if (bytes) cdk::hip_check(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c.stream), "D2H");
}
void sync(cdk::Context& c) { cdk::hip_check(hipStreamSynchronize(c.stream), "stream sync"); }

// Per-pose rotation + AABB + window on the device via the search's prep kernel.
void device_prep(cdk::Context& c, int nq, int natom, const float* quats, const float* coords, int N, float h,
                 std::vector<float>* rot, std::vector<cdk::RotMeta>& meta) {
  require(natom > 0 && natom <= cdk::kMaxAtoms, "atom count must be in [1, 512]");
  std::vector<cdk::PoseSlot> slots(nq);
  for (int i = 0; i < nq; ++i) {
    std::memset(&slots[i], 0, sizeof(cdk::PoseSlot));
    for (int k = 0; k < 4; ++k) slots[i].q[k] = quats[i * 4 + k];
  }
  cdk::LigandDesc d;
  std::memset(&d, 0, sizeof(d));
  d.natom = natom;
  d.nch = 1;
  Tmp<cdk::PoseSlot> ds(nq);
  Tmp<cdk::LigandDesc> dl(1);
  Tmp<float> dc((size_t)natom * 3), dr((size_t)nq * natom * 3);
  Tmp<cdk::AtomGrid> dag((size_t)nq * natom);
  Tmp<cdk::RotMeta> dm(nq);
  Tmp<unsigned long long> dk(nq);
  h2d(c, ds.p(), slots.data(), nq * sizeof(cdk::PoseSlot));
  h2d(c, dl.p(), &d, sizeof(d));
  h2d(c, dc.p(), coords, (size_t)natom * 3 * sizeof(float));
  cdk::hip_check(cdk::launch_prep(c.stream, ds.p(), nq, dl.p(), dc.p(), N, h, natom, 1 << 30, dr.p(), dag.p(), dm.p(),
                                 dk.p()),
                 "prep");
  meta.resize(nq);
  d2h(c, meta.data(), dm.p(), nq * sizeof(cdk::RotMeta));
  if (rot) {
    rot->resize((size_t)nq * natom * 3);
    d2h(c, rot->data(), dr.p(), rot->size() * sizeof(float));
  }
  sync(c);
}

}  // namespace

extern "C" {

const char* cdk_last_error(void) { return g_err.c_str(); }
int cdk_version(void) { return 10000; }

int cdk_context_create(int device, cdk_context** out) {
  return guard([&] {
    require(out != nullptr, "out is NULL");
    auto* c = new cdk_context();
    try {
      c->c = std::make_unique<cdk::Context>(device);
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

int cdk_context_destroy(cdk_context* ctx) {
  return guard([&] { delete ctx; });
}

int cdk_context_set_option(cdk_context* ctx, const char* key, int64_t v) {
  return guard([&] {
    require(ctx && key, "NULL argument");
    cdk::Context& c = *ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    const std::string k(key);
    if (k == "ws_budget_bytes") {
      require(v > 0, "budget must be positive");
      c.ws_budget = (size_t)v;
    } else if (k == "ref_batch") {
      require(v > 0 && v <= 100000, "ref_batch out of range");
      c.ref_batch = (int)v;
    } else if (k == "fused_chunk") {
      require(v >= 0 && v <= (1 << 24), "fused_chunk out of range");
      c.fused_chunk = (int)v;
    } else if (k == "mid_group") {
      require(v >= 0 && v <= 4096, "mid_group out of range");
      c.mid_group = (int)v;
    } else if (k == "mid_depth") {
      require(v >= 0 && v <= 3, "mid_depth must be in [0, 3]");
      c.mid_depth = (int)v;
    } else if (k == "mid_poses") {
      require(v == 0 || v == 1 || v == 2 || v == 4, "mid_poses must be 0, 1, 2 or 4");
      c.mid_poses = (int)v;
    } else if (k == "overlap_z") {
      require(v == 0 || v == 1, "overlap_z must be 0 or 1");
      c.overlap_z = (int)v;
    } else if (k == "atom_products") {
      require(v == cdk::kAtomF16 || v == cdk::kAtomF32 || v == cdk::kAtomF16x3,
              "atom_products must be 0 (FP16 matrix cores, two-part splits), 1 (FP32 matrix cores) or 2 (FP16 "
              "matrix cores, exact three-part splits)");
      c.atom_products = (int)v;
    } else if (k == "pipeline") {
      require(v >= -1 && v <= 2, "pipeline must be -1 (tuned), 0 (plane form), 1 (column form) or 2 (atom form)");
      c.pipeline = (int)v;
    } else if (k == "fwd_variant") {
      require(v >= -1 && v <= 1, "fwd_variant must be -1 (tuned), 0 (block) or 1 (register)");
      c.fwd_variant = (int)v;
    } else if (k == "inv_variant") {
      require(v >= -1 && v <= 1, "inv_variant must be -1 (tuned), 0 (block) or 1 (register)");
      c.inv_variant = (int)v;
    } else if (k == "use_graph") {
      c.use_graph = v != 0;
    } else if (k == "time_stages") {
      c.time_stages = v != 0;
    } else {
      throw cdk::Error(CDK_ERR_INVALID, "unknown option " + k);
    }
  });
}

int cdk_context_get_option(cdk_context* ctx, const char* key, int64_t* value) {
  return guard([&] {
    require(ctx && key && value, "NULL argument");
    cdk::Context& c = *ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    const std::string k(key);
    if (k == "ws_budget_bytes") *value = (int64_t)c.ws_budget;
    else if (k == "ref_batch") *value = c.ref_batch;
    else if (k == "fused_chunk") *value = c.fused_chunk;
    else if (k == "mid_group") *value = c.mid_group;
    else if (k == "fwd_variant") *value = c.fwd_variant;
    else if (k == "inv_variant") *value = c.inv_variant;
    else if (k == "last_inv_variant") *value = c.last_inv;
    else if (k == "use_graph") *value = c.use_graph;
    else if (k == "time_stages") *value = c.time_stages;
    else if (k == "last_fwd_variant") *value = c.last_fwd;
    else if (k == "last_mid_group") *value = c.last_group;
    else if (k == "mid_depth") *value = c.mid_depth;
    else if (k == "last_mid_depth") *value = c.last_depth;
    else if (k == "mid_poses") *value = c.mid_poses;
    else if (k == "pipeline") *value = c.pipeline;
    else if (k == "atom_products") *value = c.atom_products;
    else if (k == "overlap_z") *value = c.overlap_z;
    else if (k == "last_pipeline") *value = c.last_pipe;
    else if (k == "last_mid_poses") *value = c.last_poses;
    else if (k == "max_dyn_lds") *value = (int64_t)c.max_dyn_lds;
    else if (k == "workspace_bytes")
      *value = (int64_t)c.d_pool.cap;
    else throw cdk::Error(CDK_ERR_INVALID, "unknown option " + k);
  });
}

int cdk_context_stage_times(cdk_context* ctx, double* ms, int* launches, int reset) {
  return guard([&] {
    require(ctx != nullptr, "NULL context");
    cdk::Context& c = *ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    for (int i = 0; i < 8; ++i) {
      if (ms) ms[i] = c.stage_times.ms[i];
      if (launches) launches[i] = c.stage_times.launches[i];
    }
    if (reset) c.stage_times = cdk::StageTimes();
  });
}

int cdk_protein_create(cdk_context* ctx, int num_grids, int xdim, int ydim, int zdim, const float* mid3,
                       const float* len3, float spacing, const float* grid_radii, const float* values,
                       cdk_protein** out) {
  return guard([&] {
    require(ctx && grid_radii && values && out, "NULL argument");
    if (xdim != ydim || ydim != zdim) throw cdk::Error(CDK_ERR_UNSUPPORTED, "only cubic grids are supported");
    auto* p = new cdk_protein();
    p->p.ctx = ctx->c.get();
    try {
      std::lock_guard<std::mutex> lk(ctx->c->mu);
      cdk::hip_check(hipSetDevice(ctx->c->device), "hipSetDevice");
      cdk::protein_upload(p->p, num_grids, xdim, spacing, mid3, len3, grid_radii, values);
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

int cdk_protein_build(cdk_context* ctx, int natom, const double* xyz, const double* charge, const double* epsilon,
                      const double* radius, int num_probes, const float* probe_radii, int n, float spacing,
                      const float* mid3, const float* len3, double lj_cutoff, double elec_cutoff, double lj_cap,
                      double elec_cap, cdk_protein** out) {
  return guard([&] {
    require(ctx && probe_radii && out, "NULL argument");
    require(natom == 0 || (xyz && charge && epsilon && radius), "NULL atom arrays");
    require(lj_cutoff > 0 && elec_cutoff > 0 && lj_cap > 0 && elec_cap > 0, "cutoffs and caps must be positive");
    auto* p = new cdk_protein();
    p->p.ctx = ctx->c.get();
    try {
      std::lock_guard<std::mutex> lk(ctx->c->mu);
      cdk::hip_check(hipSetDevice(ctx->c->device), "hipSetDevice");
      cdk::protein_build(p->p, natom, xyz, charge, epsilon, radius, num_probes, probe_radii, n, spacing, mid3, len3,
                         lj_cutoff, elec_cutoff, lj_cap, elec_cap);
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

int cdk_protein_grids(cdk_protein* p, float* values) {
  return guard([&] {
    require(p && values, "NULL argument");
    cdk::Context& ctx = *p->p.ctx;
    std::lock_guard<std::mutex> lk(ctx.mu);
    cdk::hip_check(hipSetDevice(ctx.device), "hipSetDevice");
    cdk::hip_check(hipMemcpy(values, p->p.grid.p, (size_t)p->p.nch * p->p.G() * sizeof(float), hipMemcpyDeviceToHost),
                   "grids D2H");
  });
}

int cdk_protein_destroy(cdk_protein* p) {
  return guard([&] { delete p; });
}

int cdk_protein_spectrum(cdk_protein* p, int channel, float* out) {
  return guard([&] {
    require(p && out, "NULL argument");
    require(channel >= 0 && channel < p->p.nch, "channel out of range");
    cdk::Context& c = *p->p.ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    const size_t H = p->p.H();
    d2h(c, out, p->p.spec.as<float2>() + (size_t)channel * H, H * sizeof(float2));
    sync(c);
  });
}

int cdk_protein_bytes(cdk_protein* p, int64_t* bytes) {
  return guard([&] {
    require(p && bytes, "NULL argument");
    *bytes = (int64_t)(p->p.grid.cap + p->p.spec.cap + p->p.spec_lb.cap + p->p.spec_kz.cap + p->p.spec_mf.cap);
  });
}

int cdk_ligand_create(cdk_protein* p, int natom, const float* charge, const float* epsilon, const float* radius,
                      cdk_ligand** out) {
  return guard([&] {
    require(p && charge && epsilon && radius && out, "NULL argument");
    for (int a = 0; a < natom; ++a)
      require(std::isfinite(charge[a]) && std::isfinite(epsilon[a]) && epsilon[a] >= 0.f && std::isfinite(radius[a]),
              "ligand parameters must be finite with epsilon >= 0");
    auto* l = new cdk_ligand();
    try {
      l->l = cdk::make_ligand(p->p, natom, charge, epsilon, radius);
    } catch (...) {
      delete l;
      throw;
    }
    l->prot = p;
    *out = l;
  });
}

int cdk_ligand_destroy(cdk_ligand* l) {
  return guard([&] { delete l; });
}

int cdk_ligand_channels(const cdk_ligand* l, int* num_vdw_used, int* idx_vdw_used, int* atom_radius_index) {
  return guard([&] {
    require(l != nullptr, "NULL ligand");
    const cdk::Ligand& L = *l->l;
    if (num_vdw_used) *num_vdw_used = L.num_vdw_used;
    if (idx_vdw_used)
      for (int i = 0; i < L.num_vdw_used; ++i) idx_vdw_used[i] = L.idx_vdw_used[i];
    if (atom_radius_index)
      for (size_t r = 0; r < L.atoms_of_radius.size(); ++r)
        for (int a : L.atoms_of_radius[r]) atom_radius_index[a] = (int)r;
  });
}

int cdk_search(cdk_context* ctx, cdk_protein* p, const cdk_ligand* l, int nconf, const float* conf_coords, int nquat,
               const float* quats, int shared_quats, int engine, float* energy, int* index, float* mincoor) {
  return guard([&] {
    require(ctx && p && l && conf_coords && quats && energy, "NULL argument");
    require(l->prot == p, "ligand was created for another protein");
    require(nconf >= 0 && nquat >= 0, "negative count");
    require(engine == CDK_ENGINE_REFERENCE || engine == CDK_ENGINE_FUSED, "unknown engine");
    cdk::Job j;
    j.lig = l->l.get();
    j.nconf = nconf;
    j.coords = conf_coords;
    j.nquat = nquat;
    j.quats = quats;
    j.shared_quats = shared_quats != 0;
    std::vector<cdk::PoseResults> out;
    cdk::run_search(*ctx->c, p->p, {j}, (cdk::Engine)engine, out);
    const size_t n = (size_t)nconf * nquat;
    std::memcpy(energy, out[0].energy.data(), n * sizeof(float));
    if (index) std::memcpy(index, out[0].index.data(), n * 3 * sizeof(int));
    if (mincoor) std::memcpy(mincoor, out[0].mincoor.data(), n * 3 * sizeof(float));
  });
}

int cdk_rank_poses(cdk_protein* p, int natom, int nquat, const float* rotated, const float* energy, const int* index,
                   const float* mincoor, int n_lowest, int* out_rank, double* out_pose) {
  return guard([&] {
    require(p && energy && index && mincoor && out_rank, "NULL argument");
    require(natom > 0 && nquat >= 0 && n_lowest >= 0, "bad counts");
    std::vector<int> order(nquat);
    std::iota(order.begin(), order.end(), 0);
    // Energy order, ties by rotation index (a stable restatement of
    // sort_index, GeneDiverseConformations.h:9-20).
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return energy[a] < energy[b]; });
    const int n = std::min(n_lowest, nquat);
    const cdk::Protein& P = p->p;
    for (int i = 0; i < n; ++i) {
      const int r = order[i];
      out_rank[i] = r;
      if (!out_pose || !rotated) continue;
      for (int k = 0; k < 3; ++k) {
        // NoMini:483-486: gridMin - mincoor + idx * spacing (float product).
        const double shift = P.grid_min(k) - (double)mincoor[r * 3 + k] + (double)((float)index[r * 3 + k] * P.h);
        for (int a = 0; a < natom; ++a)
          out_pose[((size_t)i * natom + a) * 3 + k] = (double)rotated[((size_t)r * natom + a) * 3 + k] + shift;
      }
    }
  });
}

int cdk_rank_poses_quats(cdk_protein* p, int natom, int nquat, const float* conformer, const float* quats,
                         const float* energy, const int* index, const float* mincoor, int n_lowest, int* out_rank,
                         double* out_pose) {
  return guard([&] {
    require(p && conformer && quats && energy && index && mincoor && out_rank, "NULL argument");
    require(natom > 0 && nquat >= 0 && n_lowest >= 0, "bad counts");
    // rotate only the kept rotations (Rotate.cpp:4-27, bit-identical to the search's)
    std::vector<int> order(nquat);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return energy[a] < energy[b]; });
    const int n = std::min(n_lowest, nquat);
    const cdk::Protein& P = p->p;
    std::vector<float> rot((size_t)natom * 3);
    for (int i = 0; i < n; ++i) {
      const int r = order[i];
      out_rank[i] = r;
      if (!out_pose) continue;
      cdk::rotate_host(1, natom, quats + (size_t)r * 4, conformer, rot.data());
      for (int k = 0; k < 3; ++k) {
        const double shift = P.grid_min(k) - (double)mincoor[r * 3 + k] + (double)((float)index[r * 3 + k] * P.h);
        for (int a = 0; a < natom; ++a) out_pose[((size_t)i * natom + a) * 3 + k] = (double)rot[a * 3 + k] + shift;
      }
    }
  });
}

int cdk_rotate_host(int nq, int natom, const float* quats, const float* coords, float* out) {
  return guard([&] {
    require(nq >= 0 && natom >= 0, "bad counts");
    require((nq == 0 || natom == 0) || (quats && coords && out), "NULL argument");
    cdk::rotate_host(nq, natom, quats, coords, out);
  });
}

int cdk_min_max_coors(int nq, int natom, const float* coors, float* mincoors, float* maxcoors) {
  return guard([&] {
    require(nq >= 0 && natom > 0, "bad counts");
    require(nq == 0 || coors, "NULL argument");
    cdk::min_max_coors(nq, natom, coors, mincoors, maxcoors);
  });
}

int cdk_rescore_poses(cdk_protein* p, const cdk_ligand* l, int npose, const double* coords, double* energy,
                      double* grad, int* n_inside) {
  return guard([&] {
    require(p && l && energy && (coords || npose == 0), "NULL argument");
    require(l->prot == p, "ligand was set up against another protein");
    require(npose >= 0, "bad pose count");
    cdk::rescore_poses(*p->p.ctx, p->p, *l->l, npose, coords, energy, grad, n_inside);
  });
}

int cdk_refine_poses(cdk_protein* p, const cdk_ligand* l, int npose, double* coords, int steps,
                     double off_grid_penalty, double* energy) {
  return guard([&] {
    require(p && l && energy && (coords || npose == 0), "NULL argument");
    require(l->prot == p, "ligand was set up against another protein");
    require(npose >= 0 && steps >= 0 && steps <= 100000, "bad pose or step count");
    require(std::isfinite(off_grid_penalty) && off_grid_penalty >= 0, "penalty must be finite and >= 0");
    cdk::refine_poses(*p->p.ctx, p->p, *l->l, npose, coords, steps, off_grid_penalty, energy);
  });
}

int cdk_write_poses(const char* dir, int conformer, int n, int natom, const double* poses, const float* energy,
                    const char* const* atom_names, const char* const* elements, const char* energy_path) {
  return guard([&] {
    require(n >= 0 && natom > 0 && conformer >= 0, "bad counts");
    require(n == 0 || (poses && energy), "NULL argument");
    cdk::write_poses(dir, conformer, n, natom, poses, energy, atom_names, elements, energy_path);
  });
}

int cdk_grid_for_openmm(cdk_protein* p, int channel, double* values, double* bounds_nm) {
  return guard([&] {
    require(p != nullptr, "NULL protein");
    require(channel >= 0 && channel < p->p.nch, "channel out of range");
    cdk::grid_for_openmm(*p->p.ctx, p->p, channel, values, bounds_nm);
  });
}

int cdk_generate_conformers(cdk_context* ctx, int natom, const double* coords, int nrotor, const int* rotors,
                            const uint8_t* moving, const uint8_t* exclude, int ntrial, uint64_t seed,
                            double clash_dist, int max_keep, double* out, int* trial_index, int* n_kept) {
  return guard([&] {
    require(ctx && coords && out && n_kept && exclude, "NULL argument");
    require(natom > 0 && natom <= cdk::kMaxAtoms && nrotor >= 0 && ntrial >= 0 && max_keep >= 0, "bad counts");
    require(nrotor == 0 || (rotors && moving), "NULL rotor arrays");
    require(ntrial <= (1 << 20), "at most 2^20 trials per call");
    *n_kept = cdk::generate_conformers(*ctx->c, natom, coords, nrotor, rotors, moving, exclude, ntrial, seed,
                                       clash_dist, max_keep, out, trial_index);
  });
}

int cdk_batch_create(cdk_context* ctx, cdk_protein* p, cdk_batch** out) {
  return guard([&] {
    require(ctx && p && out, "NULL argument");
    auto* b = new cdk_batch();
    b->b.ctx = ctx->c.get();
    b->b.prot = &p->p;
    b->prot = p;
    *out = b;
  });
}

int cdk_batch_add(cdk_batch* b, const cdk_ligand* l, int nconf, const float* conf_coords, int nquat,
                  const float* quats, int shared_quats, int* job_id) {
  return guard([&] {
    require(b && l && conf_coords && quats, "NULL argument");
    require(l->prot == b->prot, "ligand was created for another protein");
    require(nconf >= 0 && nquat >= 0, "negative count");
    cdk::Batch::Item it;
    it.lig = l->l.get();
    it.nconf = nconf;
    it.nquat = nquat;
    it.shared = shared_quats != 0;
    it.coords.assign(conf_coords, conf_coords + (size_t)nconf * l->l->natom * 3);
    it.quats.assign(quats, quats + (size_t)(it.shared ? 1 : nconf) * nquat * 4);
    b->b.items.push_back(std::move(it));
    b->b.dirty = true;
    if (job_id) *job_id = (int)b->b.items.size() - 1;
  });
}

int cdk_batch_run(cdk_batch* b, int engine) {
  return guard([&] {
    require(b != nullptr, "NULL batch");
    require(engine == CDK_ENGINE_REFERENCE || engine == CDK_ENGINE_FUSED, "unknown engine");
    b->b.run((cdk::Engine)engine);
  });
}

int cdk_batch_poses(cdk_batch* b, int64_t* total) {
  return guard([&] {
    require(b && total, "NULL argument");
    int64_t n = 0;
    for (auto& it : b->b.items) n += (int64_t)it.nconf * it.nquat;
    *total = n;
  });
}

int cdk_batch_results(cdk_batch* b, int job_id, float* energy, int* index, float* mincoor) {
  return guard([&] {
    require(b != nullptr, "NULL batch");
    require(job_id >= 0 && job_id < (int)b->b.results().size(), "job id out of range or batch not run");
    const cdk::PoseResults& r = b->b.results()[job_id];
    if (energy) std::memcpy(energy, r.energy.data(), r.energy.size() * sizeof(float));
    if (index) std::memcpy(index, r.index.data(), r.index.size() * sizeof(int));
    if (mincoor) std::memcpy(mincoor, r.mincoor.data(), r.mincoor.size() * sizeof(float));
  });
}

int cdk_batch_geometry(cdk_batch* b, int job_id, int* box, int* windows) {
  return guard([&] {
    require(b != nullptr, "NULL batch");
    require(job_id >= 0 && job_id < (int)b->b.results().size(), "job id out of range or batch not run");
    const cdk::PoseResults& r = b->b.results()[job_id];
    if (box) std::memcpy(box, r.box.data(), r.box.size() * sizeof(int));
    if (windows) std::memcpy(windows, r.win.data(), r.win.size() * sizeof(int));
  });
}

int cdk_batch_destroy(cdk_batch* b) {
  return guard([&] { delete b; });
}

int cdk_server_create(cdk_context* ctx, int nprot, cdk_protein* const* prots, int64_t budget, int max_batch_poses,
                      cdk_server** out) {
  return guard([&] {
    require(ctx && prots && out && nprot > 0, "bad arguments");
    require(budget > 0 && max_batch_poses > 0, "budget and batch must be positive");
    std::vector<cdk::Protein*> ps;
    for (int i = 0; i < nprot; ++i) {
      require(prots[i] != nullptr && prots[i]->p.ctx == ctx->c.get(), "protein of another context");
      ps.push_back(&prots[i]->p);
    }
    auto* s = new cdk_server();
    s->s = std::make_unique<cdk::Server>(*ctx->c, ps, (size_t)budget, max_batch_poses);
    *out = s;
  });
}

int cdk_server_submit(cdk_server* s, int protein, int natom, const float* charge, const float* epsilon,
                      const float* radius, int nconf, const float* conf_coords, int nquat, const float* quats,
                      int shared_quats, int64_t* request_id) {
  return guard([&] {
    require(s && charge && epsilon && radius && conf_coords && quats && request_id, "NULL argument");
    *request_id = s->s->submit(protein, natom, charge, epsilon, radius, nconf, conf_coords, nquat, quats,
                               shared_quats != 0);
  });
}

int cdk_server_result(cdk_server* s, int64_t request_id, int wait_ms, float* energy, int* index, float* mincoor,
                      double* latency_s) {
  int rc = CDK_OK;
  const int g = guard([&] {
    require(s != nullptr, "NULL server");
    std::string err;
    rc = s->s->result(request_id, wait_ms, energy, index, mincoor, latency_s, &err);
    if (rc != CDK_OK && rc != CDK_ERR_PENDING) g_err = err;
  });
  return g != CDK_OK ? g : rc;
}

int cdk_server_stats(cdk_server* s, int n, double* stats) {
  return guard([&] {
    require(s && (stats || n <= 0), "NULL argument");
    s->s->stats(n, stats);
  });
}

int cdk_server_destroy(cdk_server* s) {
  return guard([&] { delete s; });
}

int cdk_rotate(cdk_context* ctx, int nq, int natom, const float* quats, const float* coords, float* out) {
  return guard([&] {
    require(ctx && quats && coords && out, "NULL argument");
    require(nq >= 0 && natom >= 0, "negative count");
    cdk::Context& c = *ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    Tmp<float> dq((size_t)nq * 4), dc((size_t)natom * 3), dout((size_t)nq * natom * 3);
    h2d(c, dq.p(), quats, (size_t)nq * 4 * sizeof(float));
    h2d(c, dc.p(), coords, (size_t)natom * 3 * sizeof(float));
    cdk::hip_check(cdk::launch_rotate(c.stream, nq, natom, dq.p(), dc.p(), dout.p()), "rotate");
    d2h(c, out, dout.p(), (size_t)nq * natom * 3 * sizeof(float));
    sync(c);
  });
}

int cdk_rotate_aabb(cdk_context* ctx, int nq, int natom, const float* quats, const float* coords, int grid_edge,
                    float spacing, float* rotated, float* mincoors, float* lengths, int* windows) {
  return guard([&] {
    require(ctx && quats && coords, "NULL argument");
    require(nq >= 0, "negative count");
    require(spacing > 0.f, "spacing must be positive");
    cdk::Context& c = *ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    std::vector<float> rot;
    std::vector<cdk::RotMeta> meta;
    if (nq == 0) return;
    device_prep(c, nq, natom, quats, coords, grid_edge, spacing, rotated ? &rot : nullptr, meta);
    if (rotated) std::memcpy(rotated, rot.data(), rot.size() * sizeof(float));
    for (int i = 0; i < nq; ++i)
      for (int k = 0; k < 3; ++k) {
        if (mincoors) mincoors[i * 3 + k] = meta[i].minc[k];
        if (lengths) lengths[i * 3 + k] = meta[i].len[k];
        if (windows) windows[i * 3 + k] = meta[i].win[k];
      }
  });
}

int cdk_fill_ligand_grid(cdk_context* ctx, int nq, int natom, const float* coors, const float* mincoors,
                         const float* charge, const float* epsilon, int num_vdw_used, const int* idx_vdw_used,
                         int num_vdw_grids, const int* radius_offsets, const int* radius_atoms, int xdim, int ydim,
                         int zdim, float spacing, float* values) {
  return guard([&] {
    require(ctx && coors && mincoors && charge && epsilon && radius_offsets && values, "NULL argument");
    require(nq >= 0 && natom > 0 && natom <= cdk::kMaxAtoms, "bad counts");
    require(num_vdw_used >= 0 && num_vdw_used + 1 <= cdk::kMaxCh, "bad channel count");
    require(xdim > 0 && ydim > 0 && zdim > 0 && spacing > 0.f, "bad grid");
    const int C = num_vdw_used + 1;
    std::vector<int> ch_start(C + 1), chan;
    for (int c = 0; c < num_vdw_used; ++c) {
      const int r = idx_vdw_used[c];
      require(r >= 0 && r < num_vdw_grids, "idx_vdw_used out of range");
      ch_start[c] = (int)chan.size();
      for (int i = radius_offsets[r]; i < radius_offsets[r + 1]; ++i) {
        require(radius_atoms[i] >= 0 && radius_atoms[i] < natom, "atom index out of range");
        chan.push_back(radius_atoms[i]);
      }
    }
    ch_start[num_vdw_used] = (int)chan.size();
    for (int a = 0; a < natom; ++a) chan.push_back(a);
    ch_start[C] = (int)chan.size();
    require((int)chan.size() - natom <= cdk::kMaxAtoms * 2, "channel lists too long");
    std::vector<float> wv(natom), wq(natom);
    for (int a = 0; a < natom; ++a) {
      wv[a] = std::sqrt(epsilon[a]);
      wq[a] = charge[a];
    }
    for (int c = 0; c < C; ++c) require(ch_start[c + 1] - ch_start[c] <= cdk::kMaxAtoms, "channel too long");
    cdk::Context& c = *ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    const size_t G = (size_t)xdim * ydim * zdim;
    Tmp<float> dco((size_t)nq * natom * 3), dmin((size_t)nq * 3), dwv(natom), dwq(natom), dgrid((size_t)nq * C * G);
    Tmp<int> dcs(C + 1), dch(chan.size());
    Tmp<cdk::AtomGrid> dag((size_t)nq * natom);
    h2d(c, dco.p(), coors, (size_t)nq * natom * 3 * sizeof(float));
    h2d(c, dmin.p(), mincoors, (size_t)nq * 3 * sizeof(float));
    h2d(c, dwv.p(), wv.data(), natom * sizeof(float));
    h2d(c, dwq.p(), wq.data(), natom * sizeof(float));
    h2d(c, dcs.p(), ch_start.data(), (C + 1) * sizeof(int));
    h2d(c, dch.p(), chan.data(), chan.size() * sizeof(int));
    h2d(c, dgrid.p(), values, (size_t)nq * C * G * sizeof(float));
    cdk::hip_check(cdk::launch_place(c.stream, nq, natom, dco.p(), dmin.p(), spacing, dag.p()), "place");
    cdk::hip_check(cdk::launch_fill_generic(c.stream, nq, natom, dag.p(), dcs.p(), dch.p(), dwv.p(), dwq.p(), C, xdim,
                                            ydim, zdim, dgrid.p()),
                   "fill");
    d2h(c, values, dgrid.p(), (size_t)nq * C * G * sizeof(float));
    sync(c);
  });
}

int cdk_fft_r2c_3d(cdk_context* ctx, int n, int batch, const float* in, float* out) {
  return guard([&] {
    require(ctx && in && out && batch >= 0, "bad arguments");
    if (!cdk::grid_edge_supported(n)) throw cdk::Error(CDK_ERR_UNSUPPORTED, "unsupported transform length");
    cdk::Context& c = *ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    const size_t G = (size_t)n * n * n, H = (size_t)n * n * (n / 2 + 1);
    Tmp<float> di((size_t)batch * G);
    Tmp<float2> dout((size_t)batch * H);
    h2d(c, di.p(), in, (size_t)batch * G * sizeof(float));
    cdk::hip_check(cdk::launch_r2c_3d(c.stream, n, di.p(), dout.p(), batch, c.twiddle(n)), "R2C");
    d2h(c, out, dout.p(), (size_t)batch * H * sizeof(float2));
    sync(c);
  });
}

int cdk_fft_c2r_3d(cdk_context* ctx, int n, int batch, const float* in, float* out) {
  return guard([&] {
    require(ctx && in && out && batch >= 0, "bad arguments");
    if (!cdk::grid_edge_supported(n)) throw cdk::Error(CDK_ERR_UNSUPPORTED, "unsupported transform length");
    cdk::Context& c = *ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    const size_t G = (size_t)n * n * n, H = (size_t)n * n * (n / 2 + 1);
    Tmp<float2> di((size_t)batch * H);
    Tmp<float> dout((size_t)batch * G);
    h2d(c, di.p(), in, (size_t)batch * H * sizeof(float2));
    cdk::hip_check(cdk::launch_c2r_3d(c.stream, n, di.p(), dout.p(), batch, c.twiddle(n)), "C2R");
    d2h(c, out, dout.p(), (size_t)batch * G * sizeof(float));
    sync(c);
  });
}

int cdk_conj_mult(cdk_context* ctx, int nq, int nch, int64_t H, const float* prot, float* lig) {
  return guard([&] {
    require(ctx && prot && lig && nq >= 0 && nch > 0 && nch <= cdk::kMaxCh && H > 0, "bad arguments");
    cdk::Context& c = *ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    std::vector<int> pmap(nch);
    std::iota(pmap.begin(), pmap.end(), 0);
    Tmp<float2> dp((size_t)nch * H), dl((size_t)nq * nch * H);
    Tmp<int> dm(nch);
    h2d(c, dp.p(), prot, (size_t)nch * H * sizeof(float2));
    h2d(c, dl.p(), lig, (size_t)nq * nch * H * sizeof(float2));
    h2d(c, dm.p(), pmap.data(), nch * sizeof(int));
    cdk::hip_check(cdk::launch_conj_mult(c.stream, dl.p(), dp.p(), dm.p(), nch, H, nq), "ConjMult");
    d2h(c, lig, dl.p(), (size_t)nq * nch * H * sizeof(float2));
    sync(c);
  });
}

int cdk_sum_grids(cdk_context* ctx, int nq, int nch, int64_t H, int64_t G, const float* lig, float* sum) {
  return guard([&] {
    require(ctx && lig && sum && nq >= 0 && nch > 0 && H > 0 && G > 0, "bad arguments");
    cdk::Context& c = *ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    Tmp<float2> dl((size_t)nq * nch * H), ds((size_t)nq * H);
    h2d(c, dl.p(), lig, (size_t)nq * nch * H * sizeof(float2));
    cdk::hip_check(cdk::launch_sum_grids(c.stream, dl.p(), ds.p(), nch, H, nq, (float)G), "SumGrids");
    d2h(c, sum, ds.p(), (size_t)nq * H * sizeof(float2));
    sync(c);
  });
}

int cdk_argmin_window(cdk_context* ctx, int nq, int n, const float* grids, const int* windows, float* energy,
                      int* index) {
  return guard([&] {
    require(ctx && grids && windows && energy && index && nq >= 0 && n > 0, "bad arguments");
    cdk::Context& c = *ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    const size_t G = (size_t)n * n * n;
    std::vector<cdk::RotMeta> meta(nq);
    for (int q = 0; q < nq; ++q) {
      std::memset(&meta[q], 0, sizeof(cdk::RotMeta));
      int ok = 1;
      for (int k = 0; k < 3; ++k) {
        meta[q].win[k] = std::min(windows[q * 3 + k], n);
        ok = ok && meta[q].win[k] > 0;
      }
      meta[q].valid = ok;
    }
    std::vector<unsigned long long> keys(nq, ~0ull);
    Tmp<float> dg((size_t)nq * G);
    Tmp<cdk::RotMeta> dm(nq);
    Tmp<unsigned long long> dk(nq);
    h2d(c, dg.p(), grids, (size_t)nq * G * sizeof(float));
    h2d(c, dm.p(), meta.data(), nq * sizeof(cdk::RotMeta));
    h2d(c, dk.p(), keys.data(), nq * sizeof(unsigned long long));
    cdk::hip_check(cdk::launch_argmin_full(c.stream, dg.p(), dm.p(), dk.p(), n, nq), "argmin");
    d2h(c, keys.data(), dk.p(), nq * sizeof(unsigned long long));
    sync(c);
    for (int q = 0; q < nq; ++q) {
      if (keys[q] == ~0ull) {
        energy[q] = std::numeric_limits<float>::infinity();
        index[q * 3] = index[q * 3 + 1] = index[q * 3 + 2] = -1;
      } else {
        const unsigned lin = cdk::key_index(keys[q]);
        energy[q] = cdk::key_energy(keys[q]);
        index[q * 3 + 0] = (int)(lin / (unsigned)(n * n));
        index[q * 3 + 1] = (int)((lin / (unsigned)n) % (unsigned)n);
        index[q * 3 + 2] = (int)(lin % (unsigned)n);
      }
    }
  });
}

int cdk_argmin_window_noalloc(cdk_context* ctx, int nq, int n, const float* grids, const int* windows,
                              float* energy, int* index, void* stream) {
  return guard([&] {
    require(ctx != nullptr && n > 0 && nq >= 0, "bad arguments");
    require(nq == 0 || (grids && windows && energy && index), "NULL buffer");
    if (nq > cdk::kNoAllocMaxPoses)
      throw cdk::Error(CDK_ERR_INVALID, "cdk_argmin_window_noalloc: nq exceeds the fixed maximum of " +
                                            std::to_string(cdk::kNoAllocMaxPoses) + " poses per call");
    cdk::Context& c = *ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    cdk::hip_check(hipSetDevice(c.device), "hipSetDevice");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c.stream;
    cdk::hip_check(cdk::argmin_window_noalloc(c.device, st, nq, n, grids, windows, energy, index), "argmin");
    if (!stream) cdk::hip_check(hipStreamSynchronize(st), "argmin sync");
  });
}

int cdk_rng_create(uint64_t seed, cdk_rng** out) {
  return guard([&] {
    require(out != nullptr, "NULL out");
    auto* r = new cdk_rng();
    r->g.seed(seed);
    *out = r;
  });
}

int cdk_rng_destroy(cdk_rng* r) {
  return guard([&] { delete r; });
}

int cdk_quaternion_uniform_sampling_mt(void* mt19937_64, int n, float* out) {
  return guard([&] {
    require(mt19937_64 && out && n >= 0, "bad arguments");
    cdk::quaternion_uniform_sampling(*static_cast<std::mt19937_64*>(mt19937_64), n, out);
  });
}

int cdk_quaternion_uniform_sampling(cdk_rng* r, int n, float* out) {
  return guard([&] {
    require(r && out && n >= 0, "bad arguments");
    cdk::quaternion_uniform_sampling(r->g, n, out);
  });
}

int cdk_filter_quaternions(cdk_context* ctx, const float* coor, int natom, int nq, const float* quats, float len_x,
                           float len_y, float len_z, int max_used, uint64_t shuffle_seed, int* n_used, float* out) {
  return guard([&] {
    require(ctx && coor && quats && n_used && out, "NULL argument");
    require(nq >= 0 && max_used >= 0, "negative count");
    cdk::Context& c = *ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    std::vector<int> kept;
    if (nq > 0) {
      std::vector<cdk::RotMeta> meta;
      device_prep(c, nq, natom, quats, coor, 0, 1.0f, nullptr, meta);
      // FilterQuaternions.cpp:45-53: strict '<' on every axis.
      for (int i = 0; i < nq; ++i)
        if (meta[i].len[0] < len_x && meta[i].len[1] < len_y && meta[i].len[2] < len_z) kept.push_back(i);
    }
    if ((int)kept.size() > max_used) {
      cdk::Xoshiro rng(shuffle_seed);
      for (int i = (int)kept.size() - 1; i > 0; --i) std::swap(kept[i], kept[(size_t)(rng.uniform() * (i + 1))]);
      kept.resize(max_used);
    }
    *n_used = (int)kept.size();
    for (size_t i = 0; i < kept.size(); ++i)
      for (int k = 0; k < 4; ++k) out[i * 4 + k] = quats[kept[i] * 4 + k];
  });
}

int cdk_idx_of_atoms_for_vdw_radius(int natom, const float* atom_radii, int num_vdw_grids, const float* grid_radii,
                                    int* num_vdw_used, int* idx_vdw_used, int* atom_radius_index) {
  return guard([&] {
    require(atom_radii && grid_radii && num_vdw_used && idx_vdw_used && atom_radius_index, "NULL argument");
    require(natom >= 0 && num_vdw_grids > 0, "bad counts");
    std::vector<int> count(num_vdw_grids, 0);
    for (int i = 0; i < natom; ++i) {
      int idx = 0;
      double diff = std::abs(atom_radii[i] - grid_radii[0]);
      for (int j = 1; j < num_vdw_grids; ++j)
        if (std::abs(atom_radii[i] - grid_radii[j]) < diff) {
          idx = j;
          diff = std::abs(atom_radii[i] - grid_radii[j]);
        }
      atom_radius_index[i] = idx;
      ++count[idx];
    }
    int u = 0;
    for (int j = 0; j < num_vdw_grids; ++j)
      if (count[j] > 0) idx_vdw_used[u++] = j;
    *num_vdw_used = u;
  });
}

int cdk_synth_protein_atoms(uint64_t seed, int natom, float box, double* xyz, double* charge, double* epsilon,
                            double* radius) {
  return guard([&] {
    require(xyz && charge && epsilon && radius, "NULL argument");
    require(natom > 0, "bad atom count");
    cdk::synth_protein_atoms(seed, natom, box, xyz, charge, epsilon, radius);
  });
}

int cdk_synth_protein(uint64_t seed, int natom, float box, int n, float spacing, float* values, float* radii,
                      float* mid3, float* len3) {
  return guard([&] {
    require(values && radii && mid3 && len3, "NULL argument");
    require(natom > 0 && box > 0.f && n > 1 && spacing > 0.f, "bad arguments");
    cdk::synth_protein(seed, natom, box, n, spacing, values, radii, mid3, len3);
  });
}

int cdk_synth_ligand(uint64_t seed, int natom, float* coords, float* charge, float* epsilon, float* radius) {
  return guard([&] {
    require(coords && charge && epsilon && radius && natom > 0, "bad arguments");
    cdk::synth_ligand(seed, natom, coords, charge, epsilon, radius);
  });
}

int cdk_synth_conformers(uint64_t seed, int natom, const float* coords, int nconf, float* out) {
  return guard([&] {
    require(coords && out && natom > 0 && nconf >= 0, "bad arguments");
    cdk::synth_conformers(seed, natom, coords, nconf, out);
  });
}

int cdk_read_grid_header(const char* path, int* num_grids, int* dims3, float* mid3, float* len3, float* spacing,
                         float* restraint_k) {
  return guard([&] {
    require(path != nullptr, "NULL path");
    cdk::GridHeader g;
    cdk::read_grid_header(path, g);
    if (num_grids) *num_grids = g.num_grids;
    for (int k = 0; k < 3; ++k) {
      if (dims3) dims3[k] = g.dims[k];
      if (mid3) mid3[k] = g.mid[k];
      if (len3) len3[k] = g.len[k];
    }
    if (spacing) *spacing = g.spacing;
    if (restraint_k) *restraint_k = g.restraint_k;
  });
}

int cdk_read_grids(const char* path, float* radii, float* values) {
  return guard([&] {
    require(path && radii && values, "NULL argument");
    cdk::GridHeader g;
    std::vector<float> v;
    cdk::read_grids(path, g, v);
    std::memcpy(radii, g.radii.data(), g.radii.size() * sizeof(float));
    std::memcpy(values, v.data(), v.size() * sizeof(float));
  });
}

int cdk_write_grids(const char* path, int num_grids, const int* dims3, const float* mid3, const float* len3,
                    float spacing, float restraint_k, const float* radii, const float* values) {
  return guard([&] {
    require(path && dims3 && mid3 && len3 && radii && values && num_grids >= 2, "bad arguments");
    cdk::GridHeader g;
    g.num_grids = num_grids;
    for (int k = 0; k < 3; ++k) {
      g.dims[k] = dims3[k];
      g.mid[k] = mid3[k];
      g.len[k] = len3[k];
    }
    g.spacing = spacing;
    g.restraint_k = restraint_k;
    g.radii.assign(radii, radii + num_grids - 1);
    cdk::write_grids(path, g, values);
  });
}

}  // extern "C"
