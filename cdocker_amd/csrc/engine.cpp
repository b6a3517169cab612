// Host engine: device context, resident protein spectra, ligand tables and
// the two search pipelines.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "conformers.hpp"
#include "gridbuild.hpp"
#include "kernels.hpp"
#include "noalloc.hpp"
#include "rescore.hpp"

namespace cdk {

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  const std::string msg = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) throw Error(kErrNoMem, msg);
  throw Error(kErrHip, msg);
}

DevBuf::~DevBuf() { release(); }

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

void DevBuf::ensure(size_t bytes) {
  if (bytes <= cap) return;
  release();
  if (bytes == 0) return;
  hip_check(hipMalloc(&p, bytes), "hipMalloc");
  cap = bytes;
}

Context::Context(int dev) : device(dev) {
  int n = 0;
  hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
  if (dev < 0 || dev >= n) throw Error(kErrInvalid, "device index out of range");
  hip_check(hipSetDevice(dev), "hipSetDevice");
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipStreamCreateWithFlags(&stream_z, hipStreamNonBlocking), "hipStreamCreate");
  int lds = 0;
  hip_check(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "hipDeviceGetAttribute");
  max_dyn_lds = (size_t)lds;
  hip_check(configure_fwd_reg(max_dyn_lds), "configure forward kernels");
  // the device address of this device's fixed no-alloc array, once per device
  hip_check(noalloc_init(dev), "no-alloc scratch address");
}

Context::~Context() {
  (void)hipSetDevice(device);
  if (stream) (void)hipStreamSynchronize(stream);
  if (stream_z) (void)hipStreamSynchronize(stream_z);
  for (auto& kv : twiddles) delete kv.second;
  if (h_keys) (void)hipHostFree(h_keys);
  if (h_meta) (void)hipHostFree(h_meta);
  for (hipEvent_t e : set_events) (void)hipEventDestroy(e);
  if (stream_z) (void)hipStreamDestroy(stream_z);
  if (stream) (void)hipStreamDestroy(stream);
}

const float2* Context::twiddle(int N) {
  auto it = twiddles.find(N);
  if (it != twiddles.end()) return it->second->as<float2>();
  std::vector<float2> t(N);
  const double two_pi = 6.283185307179586476925286766559;
  for (int m = 0; m < N; ++m) {
    const double a = two_pi * (double)m / (double)N;
    t[m] = make_float2((float)std::cos(a), (float)-std::sin(a));
  }
  auto* b = new DevBuf();
  b->ensure(N * sizeof(float2));
  hip_check(hipMemcpy(b->p, t.data(), N * sizeof(float2), hipMemcpyHostToDevice), "twiddle upload");
  twiddles[N] = b;
  return b->as<float2>();
}

void Context::ensure_host(size_t nslots) {
  if (nslots <= h_cap) return;
  if (h_keys) (void)hipHostFree(h_keys);
  if (h_meta) (void)hipHostFree(h_meta);
  h_keys = h_meta = nullptr;
  hip_check(hipHostMalloc(&h_keys, nslots * sizeof(unsigned long long)), "hipHostMalloc");
  hip_check(hipHostMalloc(&h_meta, nslots * sizeof(RotMeta)), "hipHostMalloc");
  h_cap = nslots;
}

namespace {

template <class T>
void upload(Context& ctx, DevBuf& b, const std::vector<T>& v) {
  b.ensure(std::max<size_t>(v.size(), 1) * sizeof(T));
  if (!v.empty())
    hip_check(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx.stream), "upload");
}

void protein_setup(Protein& p, int nch, int N, float h, const float* mid, const float* len, const float* radii) {
  if (nch < 2 || nch > kMaxCh) throw Error(kErrInvalid, "number of grids must be in [2, 16]");
  if (!grid_edge_supported(N)) throw Error(kErrUnsupported, "grid edge must be one of 32, 48, 64, 96, 128");
  if (!(h > 0.f)) throw Error(kErrInvalid, "grid spacing must be positive");
  p.nch = nch;
  p.N = N;
  p.h = h;
  for (int k = 0; k < 3; ++k) {
    p.mid[k] = mid ? mid[k] : 0.f;
    p.len[k] = len ? len[k] : (float)(N - 1) * h;
  }
  p.radii.assign(radii, radii + (nch - 1));
  p.grid.ensure((size_t)nch * p.G() * sizeof(float));
}

// Forward R2C of every channel (NoMini:177-199) plus the copies the fused
// mid stages read (line-blocked; kz-plane and matrix-core blocks at grid
// edge 64).
void protein_spectra(Protein& p) {
  Context& ctx = *p.ctx;
  const size_t H = p.H();
  p.spec.ensure((size_t)p.nch * H * sizeof(float2));
  hip_check(launch_r2c_3d(ctx.stream, p.N, p.grid.as<float>(), p.spec.as<float2>(), p.nch, ctx.twiddle(p.N)),
            "protein R2C");
  p.spec_lb.ensure((size_t)p.nch * H * sizeof(float2));
  hip_check(launch_to_line_blocks(ctx.stream, p.N, p.spec.as<float2>(), p.spec_lb.as<float2>(), p.nch),
            "protein spectra line blocks");
  if (col_supported(p.N)) {
    p.spec_kz.ensure((size_t)p.nch * H * sizeof(float2));
    hip_check(launch_to_kz_blocks(ctx.stream, p.N, p.spec.as<float2>(), p.spec_kz.as<float2>(), p.nch),
              "protein spectra kz blocks");
  }
  if (atoms_supported(p.N)) {
    p.spec_mf.ensure((size_t)p.nch * H * sizeof(float2));
    hip_check(launch_to_mfma_blocks(ctx.stream, p.N, p.spec.as<float2>(), p.spec_mf.as<float2>(), p.nch),
              "protein spectra matrix-core blocks");
  }
  hip_check(hipStreamSynchronize(ctx.stream), "protein R2C sync");
}

}  // namespace

void protein_upload(Protein& p, int nch, int N, float h, const float* mid, const float* len, const float* radii,
                    const float* values) {
  protein_setup(p, nch, N, h, mid, len, radii);
  Context& ctx = *p.ctx;
  hip_check(hipMemcpyAsync(p.grid.p, values, (size_t)nch * p.G() * sizeof(float), hipMemcpyHostToDevice, ctx.stream),
            "protein upload");
  protein_spectra(p);
}

void protein_build(Protein& p, int natom, const double* xyz, const double* charge, const double* epsilon,
                   const double* radius, int nprobe, const float* probe_radii, int N, float h, const float* mid,
                   const float* len, double lj_cutoff, double elec_cutoff, double lj_cap, double elec_cap) {
  if (natom < 0 || nprobe < 1 || nprobe > grid_build_max_probes())
    throw Error(kErrInvalid, "atom count must be >= 0 and probe count in [1, 15]");
  protein_setup(p, nprobe + 1, N, h, mid, len, probe_radii);
  Context& ctx = *p.ctx;
  std::vector<double> atoms((size_t)std::max(natom, 1) * 6);
  for (int i = 0; i < natom; ++i) {
    for (int k = 0; k < 3; ++k) atoms[(size_t)i * 6 + k] = xyz[i * 3 + k];
    atoms[(size_t)i * 6 + 3] = charge[i];
    atoms[(size_t)i * 6 + 4] = radius[i];
    atoms[(size_t)i * 6 + 5] = epsilon[i];
  }
  std::vector<double> probe(probe_radii, probe_radii + nprobe);
  DevBuf d_atoms, d_probe;
  upload(ctx, d_atoms, atoms);
  upload(ctx, d_probe, probe);
  const double gmin[3] = {p.grid_min(0), p.grid_min(1), p.grid_min(2)};
  hip_check(launch_build_grids(ctx.stream, natom, d_atoms.as<double>(), nprobe, d_probe.as<double>(), N, gmin,
                               (double)h, lj_cutoff, elec_cutoff, lj_cap, elec_cap, p.grid.as<float>()),
            "build grids");
  protein_spectra(p);
}

std::unique_ptr<Ligand> make_ligand(const Protein& p, int natom, const float* charge, const float* eps,
                                    const float* radius) {
  if (natom <= 0 || natom > kMaxAtoms) throw Error(kErrInvalid, "ligand atom count must be in [1, 512]");
  auto L = std::make_unique<Ligand>();
  L->natom = natom;
  L->charge.assign(charge, charge + natom);
  L->eps.assign(eps, eps + natom);
  L->radius.assign(radius, radius + natom);
  const int nv = p.nch - 1;
  // GetIdxOfAtomsForVdwRadius.cpp:9-23: nearest probe radius, strict '<' so
  // the first of equal distances wins; float difference, double running min.
  L->atoms_of_radius.assign(nv, {});
  L->atom_channel.assign(natom, 0);
  std::vector<int> radius_of_atom(natom);
  for (int i = 0; i < natom; ++i) {
    int idx = 0;
    double diff = std::abs(radius[i] - p.radii[0]);
    for (int j = 1; j < nv; ++j) {
      if (std::abs(radius[i] - p.radii[j]) < diff) {
        idx = j;
        diff = std::abs(radius[i] - p.radii[j]);
      }
    }
    L->atoms_of_radius[idx].push_back(i);
    radius_of_atom[i] = idx;
  }
  // GetIdxOfAtomsForVdwRadius.cpp:26-34: used radii in ascending order.
  for (int j = 0; j < nv; ++j)
    if (!L->atoms_of_radius[j].empty()) L->idx_vdw_used.push_back(j);
  L->num_vdw_used = (int)L->idx_vdw_used.size();
  const int nch = L->num_vdw_used + 1;
  if (nch > kMaxCh) throw Error(kErrInvalid, "too many channels");
  LigandDesc& d = L->desc;
  std::memset(&d, 0, sizeof(d));
  d.natom = natom;
  d.nch = nch;
  d.w_off = 0;
  d.list_off = 0;
  int pos = 0;
  for (int c = 0; c < L->num_vdw_used; ++c) {
    d.ch_start[c] = pos;
    d.pmap[c] = L->idx_vdw_used[c];
    for (int a : L->atoms_of_radius[L->idx_vdw_used[c]]) {
      L->chan_atoms.push_back(a);
      L->atom_channel[a] = c;
    }
    pos = (int)L->chan_atoms.size();
  }
  // Electrostatic channel: every atom, elec protein grid is the last one
  // (TranRotaConfSearchNoMini.cu:143-146).
  d.ch_start[L->num_vdw_used] = pos;
  d.pmap[L->num_vdw_used] = p.nch - 1;
  for (int a = 0; a < natom; ++a) L->chan_atoms.push_back(a);
  d.ch_start[nch] = (int)L->chan_atoms.size();
  d.pad_off = 0;
  for (int c = 0; c < nch; ++c) {
    d.pad_start[c] = (int)L->pad_atoms.size();
    L->pad_atoms.insert(L->pad_atoms.end(), L->chan_atoms.begin() + d.ch_start[c],
                        L->chan_atoms.begin() + d.ch_start[c + 1]);
    while (L->pad_atoms.size() % 8) L->pad_atoms.push_back(-1);
  }
  d.pad_start[nch] = (int)L->pad_atoms.size();
  L->wv.resize(natom);
  L->wq.resize(natom);
  float wmax = 0.f;
  for (int a = 0; a < natom; ++a) {
    L->wv[a] = std::sqrt(eps[a]);  // float overload, as FillLigandGrid.cpp:45
    L->wq[a] = charge[a];
    if (!std::isfinite(L->wv[a]) || !std::isfinite(L->wq[a])) throw Error(kErrInvalid, "non-finite atom weight");
    wmax = std::max({wmax, std::abs(L->wv[a]), std::abs(L->wq[a])});
  }
  // wmax * 2^e in [2^12, 2^13): e = 13 - exponent of wmax (frexp: wmax = m 2^x, m in [0.5, 1))
  int ex = 0;
  if (wmax > 0.f) (void)std::frexp(wmax, &ex);
  const int e = wmax > 0.f ? std::max(-100, std::min(100, 13 - ex)) : 0;
  d.ascale = std::ldexp(1.f, e);
  d.aunscale = std::ldexp(1.f, -e);
  return L;
}

float key_energy(unsigned long long key) {
  if (key == ~0ull) return std::numeric_limits<float>::infinity();
  unsigned u = (unsigned)(key >> 32);
  u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

unsigned key_index(unsigned long long key) { return (unsigned)(key & 0xffffffffu); }

namespace {

// Upper bound of the occupied box edge of a conformer over all rotations:
// the extent along any axis is at most the largest pairwise distance.
int box_bound(const float* xyz, int natom, float h, int N, double qscale) {
  double d2 = 0;
  for (int i = 0; i < natom; ++i)
    for (int j = i + 1; j < natom; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) {
        const double d = (double)xyz[i * 3 + k] - (double)xyz[j * 3 + k];
        s += d * d;
      }
      d2 = std::max(d2, s);
    }
  const int b = (int)std::floor(std::sqrt(d2) * qscale / (double)h) + 3;
  return std::min(N, std::max(b, 2));
}

// Rotation scales lengths by |q|^2; the bound must hold for non-unit input.
double quat_scale(const float* q, size_t n) {
  double m = 1.0;
  for (size_t i = 0; i < n; ++i) {
    const double s = (double)q[i * 4] * q[i * 4] + (double)q[i * 4 + 1] * q[i * 4 + 1] +
                     (double)q[i * 4 + 2] * q[i * 4 + 2] + (double)q[i * 4 + 3] * q[i * 4 + 3];
    m = std::max(m, s);
  }
  return m * (1.0 + 1e-5);
}

void build_tables(const Protein& prot, const std::vector<Job>& jobs, Prepared& t) {
  std::map<const Ligand*, int> lig_index;
  for (const Job& j : jobs) {
    if (!j.lig) throw Error(kErrInvalid, "job without ligand");
    if (j.nconf < 0 || j.nquat < 0) throw Error(kErrInvalid, "negative conformer or quaternion count");
    if (lig_index.count(j.lig)) continue;
    const Ligand& L = *j.lig;
    LigandDesc d = L.desc;
    d.list_off = (int)t.chan.size();
    d.pad_off = (int)t.pad.size();
    d.w_off = (int)t.wv.size();
    t.chan.insert(t.chan.end(), L.chan_atoms.begin(), L.chan_atoms.end());
    t.pad.insert(t.pad.end(), L.pad_atoms.begin(), L.pad_atoms.end());
    t.wv.insert(t.wv.end(), L.wv.begin(), L.wv.end());
    t.wq.insert(t.wq.end(), L.wq.begin(), L.wq.end());
    for (int c = 0; c < d.nch; ++c)
      if (d.pmap[c] < 0 || d.pmap[c] >= prot.nch) throw Error(kErrInvalid, "ligand built for another protein");
    lig_index[j.lig] = (int)t.ligs.size();
    t.ligs.push_back(d);
    t.astride = std::max(t.astride, L.natom);
    t.cmax = std::max(t.cmax, d.nch);
    for (int c = 0; c < d.nch; ++c) t.max_chan_atoms = std::max(t.max_chan_atoms, d.ch_start[c + 1] - d.ch_start[c]);
  }
  // Padding ligand: zero atoms, so its poses are never valid.
  LigandDesc dummy;
  std::memset(&dummy, 0, sizeof(dummy));
  dummy.nch = 1;
  dummy.ascale = dummy.aunscale = 1.f;
  t.dummy_lig = (int)t.ligs.size();
  t.ligs.push_back(dummy);
  if (t.chan.empty()) t.chan.push_back(0);
  if (t.wv.empty()) {
    t.wv.push_back(0.f);
    t.wq.push_back(0.f);
  }
  size_t nslots = 0;
  for (const Job& j : jobs) nslots += j.npose();
  if (nslots > (size_t)std::numeric_limits<int>::max()) throw Error(kErrInvalid, "too many poses in one call");
  t.slots.reserve(nslots);
  for (const Job& j : jobs) {
    t.job_first.push_back(t.slots.size());
    const size_t np = j.npose();
    if (np == 0) continue;
    const int li = lig_index[j.lig];
    const int na = j.lig->natom;
    const size_t nq_total = (size_t)(j.shared_quats ? 1 : j.nconf) * j.nquat;
    for (size_t i = 0; i < nq_total * 4; ++i)
      if (!std::isfinite(j.quats[i])) throw Error(kErrInvalid, "non-finite quaternion");
    const double qscale = quat_scale(j.quats, nq_total);
    int cur_conf = -1;
    int off = 0;
    for (size_t pose = j.pose_begin; pose < j.pose_begin + np; ++pose) {
      const int c = (int)(pose / (size_t)j.nquat), r = (int)(pose % (size_t)j.nquat);
      if (c != cur_conf) {
        cur_conf = c;
        if (t.coords.size() + (size_t)na * 3 > (size_t)std::numeric_limits<int>::max())
          throw Error(kErrInvalid, "coordinate table too large");
        off = (int)t.coords.size();
        const float* xyz = j.coords + (size_t)c * na * 3;
        for (int a = 0; a < na * 3; ++a)
          if (!std::isfinite(xyz[a])) throw Error(kErrInvalid, "non-finite ligand coordinate");
        t.coords.insert(t.coords.end(), xyz, xyz + (size_t)na * 3);
        t.boxmax = std::max(t.boxmax, box_bound(xyz, na, prot.h, prot.N, qscale));
      }
      const float* q = j.shared_quats ? j.quats : j.quats + (size_t)c * j.nquat * 4;
      PoseSlot sl;
      std::memset(&sl, 0, sizeof(sl));
      for (int k = 0; k < 4; ++k) sl.q[k] = q[r * 4 + k];
      sl.lig = li;
      sl.coord_off = off;
      t.slots.push_back(sl);
    }
  }
  if (t.coords.empty()) t.coords.resize(3, 0.f);
  t.astride = std::max(t.astride, 1);
  t.cmax = std::max(t.cmax, 1);
}


struct EventTimer {
  Context& ctx;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
  explicit EventTimer(Context& c) : ctx(c) {}
  ~EventTimer() {
    for (auto& e : ev) {
      (void)hipEventDestroy(e.second.first);
      (void)hipEventDestroy(e.second.second);
    }
  }
  hipEvent_t begin(int stage) {
    if (!ctx.time_stages) return nullptr;
    hipEvent_t a, b;
    hip_check(hipEventCreate(&a), "hipEventCreate");
    hip_check(hipEventCreate(&b), "hipEventCreate");
    ev.push_back({stage, {a, b}});
    hip_check(hipEventRecord(a, ctx.stream), "hipEventRecord");
    return b;
  }
  void end(hipEvent_t b) {
    if (b) hip_check(hipEventRecord(b, ctx.stream), "hipEventRecord");
  }
  void collect() {
    for (auto& e : ev) {
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, e.second.first, e.second.second), "hipEventElapsedTime");
      ctx.stage_times.ms[e.first] += ms;
      ctx.stage_times.launches[e.first] += 1;
    }
  }
};

}  // namespace

namespace {

// Before a working array is reallocated: the context's queued work (a
// server's earlier launch set may still run on either stream) is done with
// the old memory.
void drain_streams(Context& ctx) {
  hip_check(hipStreamSynchronize(ctx.stream), "hipStreamSynchronize");
  hip_check(hipStreamSynchronize(ctx.stream_z), "hipStreamSynchronize");
}

void ensure_work(Context& ctx, size_t bytes_needed, DevBuf& b) {
  if (b.cap >= bytes_needed) return;
  drain_streams(ctx);
  b.ensure(bytes_needed);
  ++ctx.ws_generation;
}

// Carves the five launch-set arrays for `chunk` poses (bytes per pose
// per[i]: rotated atoms, placements, A, T, second T2) from the pool, each
// 256-byte aligned, growing the pool if needed.
void carve_work(Context& ctx, size_t chunk, const size_t (&per)[5]) {
  size_t off[5], end = 0;
  for (int i = 0; i < 5; ++i) {
    off[i] = end;
    end += (chunk * per[i] + 255) & ~(size_t)255;
  }
  if (ctx.d_pool.cap < end) {
    drain_streams(ctx);
    ctx.d_pool.ensure(end);
    ++ctx.ws_generation;
  }
  char* base = static_cast<char*>(ctx.d_pool.p);
  float* rot = reinterpret_cast<float*>(base + off[0]);
  AtomGrid* ag = reinterpret_cast<AtomGrid*>(base + off[1]);
  float2* A = reinterpret_cast<float2*>(base + off[2]);
  float2* T = reinterpret_cast<float2*>(base + off[3]);
  float2* T2b = per[4] ? reinterpret_cast<float2*>(base + off[4]) : nullptr;
  if (rot != ctx.w_rot || ag != ctx.w_ag || A != ctx.w_A || T != ctx.w_T || T2b != ctx.w_T2b) ++ctx.ws_generation;
  ctx.w_rot = rot;
  ctx.w_ag = ag;
  ctx.w_A = A;
  ctx.w_T = T;
  ctx.w_T2b = T2b;
}

// Reference dataflow: one launch sequence per batch of <= B rotations of one
// conformer (TranRotaConfSearchNoMini.cu:385-464).
void enqueue_reference(Context& ctx, const Protein& prot, Prepared& t, EventTimer& timer) {
  const int N = prot.N;
  const size_t G = prot.G(), H = prot.H();
  const float2* tw = ctx.twiddle(N);
  const int B = std::max(1, ctx.ref_batch);
  const int C = t.cmax;
  const size_t per[5] = {(size_t)t.astride * 3 * sizeof(float), (size_t)t.astride * sizeof(AtomGrid), 0, 0, 0};
  carve_work(ctx, (size_t)B, per);
  ensure_work(ctx, (size_t)B * C * G * sizeof(float), ctx.d_L);
  ensure_work(ctx, (size_t)B * C * H * sizeof(float2), ctx.d_Lhat);
  ensure_work(ctx, (size_t)B * H * sizeof(float2), ctx.d_S);
  ensure_work(ctx, (size_t)B * G * sizeof(float), ctx.d_E);
  const PoseSlot* d_slots = t.d_slots.as<PoseSlot>();
  const LigandDesc* d_ligs = t.d_ligs.as<LigandDesc>();
  unsigned long long* d_keys = t.d_keys.as<unsigned long long>();
  RotMeta* d_meta = t.d_meta.as<RotMeta>();
  const size_t total = t.slots.size();
  size_t s = 0;
  while (s < total) {
    const int lig_idx = t.slots[s].lig;
    const int coff = t.slots[s].coord_off;
    size_t e2 = s + 1;
    while (e2 < total && (int)(e2 - s) < B && t.slots[e2].coord_off == coff && t.slots[e2].lig == lig_idx) ++e2;
    const int n = (int)(e2 - s);
    const int nch = t.ligs[lig_idx].nch;
    const int* pmap = reinterpret_cast<const int*>(reinterpret_cast<const char*>(d_ligs + lig_idx) +
                                                   offsetof(LigandDesc, pmap));
    const PoseSlot* sl = d_slots + s;
    hipEvent_t e;
    e = timer.begin(0);
    hip_check(launch_prep(ctx.stream, sl, n, d_ligs, t.d_coords.as<float>(), N, prot.h, t.astride, N,
                          ctx.w_rot, ctx.w_ag, d_meta + s, d_keys + s),
              "prep");
    timer.end(e);
    e = timer.begin(1);
    hip_check(launch_spread_full(ctx.stream, sl, n, d_ligs, t.d_chan.as<int>(), t.d_wv.as<float>(), t.d_wq.as<float>(),
                                 ctx.w_ag, t.astride, d_meta + s, N, nch, ctx.d_L.as<float>()),
              "spread");
    timer.end(e);
    e = timer.begin(2);
    hip_check(launch_r2c_3d(ctx.stream, N, ctx.d_L.as<float>(), ctx.d_Lhat.as<float2>(), n * nch, tw), "ligand R2C");
    timer.end(e);
    e = timer.begin(3);
    hip_check(launch_conj_mult(ctx.stream, ctx.d_Lhat.as<float2>(), prot.spec.as<float2>(), pmap, nch, (long long)H,
                               n),
              "ConjMult");
    timer.end(e);
    e = timer.begin(4);
    hip_check(launch_sum_grids(ctx.stream, ctx.d_Lhat.as<float2>(), ctx.d_S.as<float2>(), nch, (long long)H, n,
                               (float)G),
              "SumGrids");
    timer.end(e);
    e = timer.begin(5);
    hip_check(launch_c2r_3d(ctx.stream, N, ctx.d_S.as<float2>(), ctx.d_E.as<float>(), n, tw), "C2R");
    timer.end(e);
    e = timer.begin(6);
    hip_check(launch_argmin_full(ctx.stream, ctx.d_E.as<float>(), d_meta + s, d_keys + s, N, n), "argmin");
    timer.end(e);
    s = e2;
  }
}

// Whether a fused search on this protein may run the atom form with its
// z-inverse stage overlapped (two T2 buffers).
bool overlap_possible(const Context& ctx, const Protein& prot) {
  return ctx.overlap_z && !ctx.time_stages && atoms_supported(prot.N) && prot.spec_mf.p != nullptr &&
         (ctx.pipeline < 0 || ctx.pipeline == kPipeAtoms);
}

// Bytes per pose of each launch-set working array: rotated atoms, atom
// placements, A (the plane and column forms' ligand spectra), T (correlation
// spectra, the atom form's T2: at grid edges 96 and 128 the window's y rows
// of the product spectra), and the overlapped atom form's second T2 buffer
// when that form may run.
void per_pose_bytes(const Context& ctx, const Protein& prot, const Prepared& t, size_t (&s)[5]) {
  const int N = prot.N, ZH = N / 2 + 1;
  const size_t plane = (size_t)N * ZH;
  s[0] = (size_t)t.astride * 3 * sizeof(float);
  s[1] = (size_t)t.astride * sizeof(AtomGrid);
  s[2] = (size_t)t.cmax * t.boxmax * plane * sizeof(float2);
  s[3] = (size_t)N * plane * sizeof(float2);
  s[4] = overlap_possible(ctx, prot) ? s[3] : 0;
}

// Poses per launch set: as many as the working arrays of the search allow
// within ws_budget (or the forced fused_chunk), at most the search's poses.
size_t fused_chunk(Context& ctx, const Protein& prot, const Prepared& t) {
  if (ctx.fused_chunk > 0) return std::min((size_t)ctx.fused_chunk, std::max<size_t>(t.slots.size(), 1));
  size_t s[5];
  per_pose_bytes(ctx, prot, t, s);
  size_t per = 0;
  for (int i = 0; i < 5; ++i) per += s[i];
  // each array's start is 256-byte aligned: 5 x 256 bytes of slack
  const size_t room = ctx.ws_budget > 5 * 256 ? ctx.ws_budget - 5 * 256 : 0;
  const size_t chunk = std::max<size_t>(1, room / std::max<size_t>(per, 1));
  return std::min(chunk, std::max<size_t>(t.slots.size(), 1));
}

// The launch-set working arrays for `chunk` poses carved from the pool
// (grown when a search needs more; the pool is never larger than the largest
// search's need, so at most ws_budget).
void ensure_fused_work(Context& ctx, const Protein& prot, const Prepared& t, size_t chunk) {
  size_t s[5];
  per_pose_bytes(ctx, prot, t, s);
  carve_work(ctx, chunk, s);
}

struct Tuned {
  int pipe = kPipePlanes;
  int fwd = kFwdBlock, inv = kInvBlock, group = 16, depth = 1, poses = 1;  // plane form
  int col_group = 4;                                                       // column form
  int atom_group = 4;                                                      // atom form
  int mid_group() const { return pipe == kPipeColumns ? col_group : pipe == kPipeAtoms ? atom_group : group; }
};

// Autotuner, on the first launch set of a search:
//  - the pipeline form (plane or column form) and, for the plane form, the
//    forward and inverse plane stage forms, once per grid size: the forms
//    round differently, so fixing them per (context, N) keeps results
//    independent of how poses are batched;
//  - the mid-stage schedule per launch shape (N, channels, box bound, chunk):
//    plane form group, ring depth and poses per protein load, column form
//    group.  Schedules only reorder work; results are bit-identical.
// Choices are made by timing every candidate twice and keeping the faster.
Tuned tune_fused(Context& ctx, const Protein& prot, Prepared& t, size_t chunk) {
  const int N = prot.N;
  Tuned tu;
  const bool fwd_reg_ok = fwd_reg_supported(N);
  const bool reg_ok = inv_reg_supported(N);
  const bool col_ok = col_supported(N) && prot.spec_kz.p != nullptr;
  const bool atoms_ok = atoms_supported(N) && prot.spec_mf.p != nullptr;
  // The plane form's schedule is kept per (N, channels, box bound, launch
  // size rounded up to a power of two): a server's launch sets vary in size
  // from set to set, and a key per exact size would re-run the tuner on
  // nearly every set.
  size_t chunk_class = 1;
  while (chunk_class < chunk) chunk_class <<= 1;
  const std::array<int, 4> gkey = {N, t.cmax, t.boxmax, (int)std::min<size_t>(chunk_class, 1 << 30)};
  // The column and atom forms' group choice is kept per (N, channels): their
  // best group hardly moves with the box bound or the launch size, and a
  // server sees a new (box bound, chunk) pair on many launch sets.
  const std::array<int, 4> ckey = {N, t.cmax, 0, 0};
  bool need_pipe = false, need_fwd = false, need_inv = false, need_group = false, need_col = false,
       need_atom = false;
  if (ctx.pipeline >= 0) {
    tu.pipe = (ctx.pipeline == kPipeColumns && col_ok)  ? kPipeColumns
              : (ctx.pipeline == kPipeAtoms && atoms_ok) ? kPipeAtoms
                                                         : kPipePlanes;
  } else {
    auto it = ctx.tuned_pipe.find(N);
    if (it != ctx.tuned_pipe.end())
      tu.pipe = it->second;
    else if (col_ok || atoms_ok)
      need_pipe = true;
    else
      ctx.tuned_pipe[N] = tu.pipe = kPipePlanes;
  }
  const bool planes = need_pipe || tu.pipe == kPipePlanes;
  const bool cols = (need_pipe && col_ok) || tu.pipe == kPipeColumns;
  const bool atoms = (need_pipe && atoms_ok) || tu.pipe == kPipeAtoms;
  if (ctx.inv_variant >= 0) {
    tu.inv = (ctx.inv_variant == kInvReg && reg_ok) ? kInvReg : kInvBlock;
  } else {
    auto it = ctx.tuned_inv.find(N);
    if (it != ctx.tuned_inv.end())
      tu.inv = it->second;
    else if (reg_ok)
      need_inv = planes;
    else
      ctx.tuned_inv[N] = tu.inv = kInvBlock;
  }
  if (ctx.fwd_variant >= 0) {
    tu.fwd = (ctx.fwd_variant == kFwdReg && fwd_reg_ok) ? kFwdReg : kFwdBlock;
  } else {
    auto it = ctx.tuned_fwd.find(N);
    if (it != ctx.tuned_fwd.end())
      tu.fwd = it->second;
    else if (fwd_reg_ok)
      need_fwd = planes;
    else
      ctx.tuned_fwd[N] = tu.fwd = kFwdBlock;
  }
  const int max_depth = mid_max_depth(N);
  const int max_poses = mid_max_poses(N);
  if (ctx.mid_group > 0) {
    tu.group = ctx.mid_group;
    tu.depth = ctx.mid_depth > 0 ? std::min(ctx.mid_depth, max_depth) : std::min(2, max_depth);
    tu.poses = ctx.mid_poses > 0 ? std::min(ctx.mid_poses, max_poses) : 1;
    tu.col_group = ctx.mid_group;
    tu.atom_group = ctx.mid_group;
  } else {
    if (planes) {
      auto it = ctx.tuned_group.find(gkey);
      if (it != ctx.tuned_group.end()) {
        tu.group = it->second[0];
        tu.depth = it->second[1];
        tu.poses = it->second[2];
      } else {
        need_group = true;
      }
    }
    if (cols) {
      auto it = ctx.tuned_col_group.find(ckey);
      if (it != ctx.tuned_col_group.end())
        tu.col_group = it->second;
      else
        need_col = true;
    }
    if (atoms) {
      auto it = ctx.tuned_atom_group.find(ckey);
      if (it != ctx.tuned_atom_group.end())
        tu.atom_group = it->second;
      else
        need_atom = true;
    }
  }
  if (!need_pipe && !need_fwd && !need_inv && !need_group && !need_col && !need_atom) return tu;

  const int n = (int)std::min(chunk, t.slots.size());
  const float2* tw = ctx.twiddle(N);
  const PoseSlot* sl = t.d_slots.as<PoseSlot>();
  const LigandDesc* d_ligs = t.d_ligs.as<LigandDesc>();
  RotMeta* d_meta = t.d_meta.as<RotMeta>();
  unsigned long long* d_keys = t.d_keys.as<unsigned long long>();
  const float invG = 1.0f / (float)prot.G();
  hip_check(launch_prep(ctx.stream, sl, n, d_ligs, t.d_coords.as<float>(), N, prot.h, t.astride, t.boxmax,
                        ctx.w_rot, ctx.w_ag, d_meta, d_keys),
            "tune prep");
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  auto timed = [&](auto&& launch) {
    float best = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {
      hip_check(hipEventRecord(e0, ctx.stream), "hipEventRecord");
      launch();
      hip_check(hipEventRecord(e1, ctx.stream), "hipEventRecord");
      hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      best = std::min(best, ms);
    }
    return best;
  };
  // The fastest of the group sizes 1, 2, 4, 8 for a mid stage: three
  // round-robin passes, each size's minimum, so clock ramps and the order of
  // the trials do not pick the size.
  auto best_group = [&](auto&& time_group, int& group) {
    const int sizes[4] = {1, 2, 4, 8};
    float t[4] = {1e30f, 1e30f, 1e30f, 1e30f};
    for (int pass = 0; pass < 3; ++pass)
      for (int i = 0; i < 4; ++i) t[i] = std::min(t[i], time_group(sizes[i]));
    int b = 0;
    for (int i = 1; i < 4; ++i)
      if (t[i] < t[b]) b = i;
    group = sizes[b];
    return t[b];
  };
  auto fwd = [&](int variant) {
    hip_check(launch_fwd_planes(ctx.stream, N, variant, sl, n, d_ligs, t.d_chan.as<int>(), t.d_wv.as<float>(),
                                t.d_wq.as<float>(), ctx.w_ag, t.astride, d_meta, t.cmax, t.boxmax,
                                t.max_chan_atoms, ctx.w_A, tw),
              "tune fwd planes");
  };
  auto mid = [&](int g, int depth, int np) {
    return timed([&] {
      hip_check(launch_mid(ctx.stream, N, depth, np, sl, d_ligs, d_meta, n, g, ctx.w_A, t.cmax,
                           t.boxmax, prot.spec_lb.as<float2>(), prot.nch, ctx.w_T, invG, tw),
                "tune mid");
    });
  };
  auto inv = [&](int variant) {
    hip_check(launch_inv_argmin(ctx.stream, N, variant, d_meta, n, N - 2, ctx.w_T, d_keys, tw),
              "tune inverse");
  };
  float plane_ms = 0.f, col_ms = 0.f, atom_ms = 0.f;
  try {
    if (need_fwd || need_inv || need_group || (need_pipe && planes)) {
      if (need_fwd) {
        const float tb = timed([&] { fwd(kFwdBlock); });
        const float twv = timed([&] { fwd(kFwdReg); });
        tu.fwd = twv < tb ? kFwdReg : kFwdBlock;
        ctx.tuned_fwd[N] = tu.fwd;
        plane_ms += std::min(tb, twv);
      } else {
        plane_ms += timed([&] { fwd(tu.fwd); });
      }
      if (need_group) {
        // One pose at a time: group size at the shallowest ring, then the
        // ring depth at that group.  Then runs of 2 and 4 poses per protein
        // load over the group sizes.
        const int d0 = std::min(2, max_depth);
        float best = 1e30f;
        for (int g : {4, 8, 16, 32}) {
          const float ms = mid(g, d0, 1);
          if (ms < best) {
            best = ms;
            tu.group = g;
          }
        }
        tu.depth = d0;
        for (int d = d0 + 1; d <= max_depth; ++d) {
          const float ms = mid(tu.group, d, 1);
          if (ms < best) {
            best = ms;
            tu.depth = d;
          }
        }
        tu.poses = 1;
        for (int np = 2; np <= max_poses; np *= 2)
          for (int g : {4, 8, 16, 32}) {
            const float ms = mid(g, tu.depth, np);
            if (ms < best) {
              best = ms;
              tu.group = g;
              tu.poses = np;
            }
          }
        if (ctx.tuned_group.size() >= kMaxTunedShapes) ctx.tuned_group.erase(ctx.tuned_group.begin());
        ctx.tuned_group[gkey] = {tu.group, tu.depth, tu.poses};
        plane_ms += best;
      } else {
        plane_ms += mid(tu.group, tu.depth, tu.poses);
      }
      if (need_inv) {
        const float tb = timed([&] { inv(kInvBlock); });
        const float tr = timed([&] { inv(kInvReg); });
        tu.inv = tr < tb ? kInvReg : kInvBlock;
        ctx.tuned_inv[N] = tu.inv;
        plane_ms += std::min(tb, tr);
      } else {
        plane_ms += timed([&] { inv(tu.inv); });
      }
    }
    if (need_col || (need_pipe && cols)) {
      col_ms += timed([&] {
        hip_check(launch_fwd_z(ctx.stream, N, sl, n, d_ligs, t.d_chan.as<int>(), t.d_wv.as<float>(),
                               t.d_wq.as<float>(), ctx.w_ag, t.astride, d_meta, t.cmax, t.boxmax,
                               t.max_chan_atoms, ctx.w_A, tw),
                  "tune column forward");
      });
      auto midc = [&](int g) {
        return timed([&] {
          hip_check(launch_mid_kz(ctx.stream, N, sl, d_ligs, d_meta, n, g, ctx.w_A, t.cmax, t.boxmax,
                                  prot.spec_kz.as<float2>(), ctx.w_T, invG, tw),
                    "tune column mid");
        });
      };
      if (need_col) {
        col_ms += best_group(midc, tu.col_group);
        ctx.tuned_col_group[ckey] = tu.col_group;
      } else {
        col_ms += midc(tu.col_group);
      }
      col_ms += timed([&] {
        hip_check(launch_inv_z(ctx.stream, N, d_meta, n, ctx.w_T, d_keys, tw), "tune column inverse");
      });
    }
    if (need_atom || (need_pipe && atoms)) {
      auto mida = [&](int g) {
        return timed([&] {
          hip_check(launch_mid_atoms(ctx.stream, N, sl, d_ligs, t.d_pad.as<int>(), t.d_wv.as<float>(),
                                     t.d_wq.as<float>(), ctx.w_ag, t.astride, d_meta, n, g,
                                     prot.spec_mf.as<float2>(), ctx.w_T, invG, tw, ctx.atom_products),
                    "tune atom mid");
        });
      };
      if (need_atom) {
        atom_ms += best_group(mida, tu.atom_group);
        ctx.tuned_atom_group[ckey] = tu.atom_group;
      } else {
        atom_ms += mida(tu.atom_group);
      }
      atom_ms += timed([&] {
        hip_check(launch_inv_z(ctx.stream, N, d_meta, n, ctx.w_T, d_keys, tw), "tune atom inverse");
      });
    }
    if (need_pipe) {
      tu.pipe = kPipePlanes;
      float best = plane_ms;
      if (cols && col_ms < best) {
        best = col_ms;
        tu.pipe = kPipeColumns;
      }
      if (atoms && atom_ms < best) {
        best = atom_ms;
        tu.pipe = kPipeAtoms;
      }
      ctx.tuned_pipe[N] = tu.pipe;
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return tu;
}

// The atom form's z-inverse stage runs on its own stream, one launch set
// behind (not when stages are timed: their events are on the launch stream).
bool overlap_z_stage(const Context& ctx, const Tuned& tu) {
  return tu.pipe == kPipeAtoms && ctx.overlap_z && !ctx.time_stages;
}

// The ordering events of an overlapped search (its second T2 buffer is
// carved with the other arrays), made before any capture.
void ensure_overlap(Context& ctx, const Prepared& t, size_t chunk) {
  if (!ctx.w_T2b) throw Error(kErrInvalid, "overlapped z-inverse stage without its second buffer");
  const size_t sets = (t.slots.size() + chunk - 1) / chunk;
  while (ctx.set_events.size() < 2 * sets) {
    hipEvent_t e = nullptr;
    hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    ctx.set_events.push_back(e);
  }
}

// Fused pipeline: four launches per chunk of poses, any mix of ligands, in
// the plane or the column form.
void enqueue_fused(Context& ctx, const Protein& prot, Prepared& t, size_t chunk, const Tuned& tu,
                   EventTimer& timer) {
  const int N = prot.N;
  const size_t G = prot.G();
  const float2* tw = ctx.twiddle(N);
  const PoseSlot* d_slots = t.d_slots.as<PoseSlot>();
  const LigandDesc* d_ligs = t.d_ligs.as<LigandDesc>();
  unsigned long long* d_keys = t.d_keys.as<unsigned long long>();
  RotMeta* d_meta = t.d_meta.as<RotMeta>();
  const size_t total = t.slots.size();
  const int group = std::max(1, tu.group);
  const bool ovl = overlap_z_stage(ctx, tu);
  size_t last_set = 0;
  for (size_t s = 0; s < total; s += chunk) {
    const int n = (int)std::min(chunk, total - s);
    const PoseSlot* sl = d_slots + s;
    const size_t si = s / chunk;  // launch set index
    hipEvent_t e;
    e = timer.begin(0);
    hip_check(launch_prep(ctx.stream, sl, n, d_ligs, t.d_coords.as<float>(), N, prot.h, t.astride, t.boxmax,
                          ctx.w_rot, ctx.w_ag, d_meta + s, d_keys + s),
              "prep");
    timer.end(e);
    if (tu.pipe == kPipeAtoms) {
      // overlapped: T2 alternates between two buffers; set si's mid stage
      // waits for set si - 2's z-inverse, the last reader of its buffer
      float2* T2 = ovl && (si & 1) ? ctx.w_T2b : ctx.w_T;
      if (ovl && si >= 2) hip_check(hipStreamWaitEvent(ctx.stream, ctx.set_events[2 * (si - 2) + 1], 0), "wait");
      e = timer.begin(2);
      hip_check(launch_mid_atoms(ctx.stream, N, sl, d_ligs, t.d_pad.as<int>(), t.d_wv.as<float>(),
                                 t.d_wq.as<float>(), ctx.w_ag, t.astride, d_meta + s, n, tu.atom_group,
                                 prot.spec_mf.as<float2>(), T2, 1.0f / (float)G, tw, ctx.atom_products),
                "atom mid");
      timer.end(e);
      if (ovl) {
        hip_check(hipEventRecord(ctx.set_events[2 * si], ctx.stream), "record");
        hip_check(hipStreamWaitEvent(ctx.stream_z, ctx.set_events[2 * si], 0), "wait");
        hip_check(launch_inv_z(ctx.stream_z, N, d_meta + s, n, T2, d_keys + s, tw), "atom inverse + argmin");
        hip_check(hipEventRecord(ctx.set_events[2 * si + 1], ctx.stream_z), "record");
        last_set = si;
      } else {
        e = timer.begin(3);
        hip_check(launch_inv_z(ctx.stream, N, d_meta + s, n, T2, d_keys + s, tw), "atom inverse + argmin");
        timer.end(e);
      }
      continue;
    }
    if (tu.pipe == kPipeColumns) {
      e = timer.begin(1);
      hip_check(launch_fwd_z(ctx.stream, N, sl, n, d_ligs, t.d_chan.as<int>(), t.d_wv.as<float>(),
                             t.d_wq.as<float>(), ctx.w_ag, t.astride, d_meta + s, t.cmax, t.boxmax,
                             t.max_chan_atoms, ctx.w_A, tw),
                "column forward");
      timer.end(e);
      e = timer.begin(2);
      hip_check(launch_mid_kz(ctx.stream, N, sl, d_ligs, d_meta + s, n, tu.col_group, ctx.w_A, t.cmax,
                              t.boxmax, prot.spec_kz.as<float2>(), ctx.w_T, 1.0f / (float)G, tw),
                "column mid");
      timer.end(e);
      e = timer.begin(3);
      hip_check(launch_inv_z(ctx.stream, N, d_meta + s, n, ctx.w_T, d_keys + s, tw),
                "column inverse + argmin");
      timer.end(e);
      continue;
    }
    e = timer.begin(1);
    hip_check(launch_fwd_planes(ctx.stream, N, tu.fwd, sl, n, d_ligs, t.d_chan.as<int>(), t.d_wv.as<float>(),
                                t.d_wq.as<float>(), ctx.w_ag, t.astride, d_meta + s, t.cmax, t.boxmax,
                                t.max_chan_atoms, ctx.w_A, tw),
              "fwd planes");
    timer.end(e);
    e = timer.begin(2);
    hip_check(launch_mid(ctx.stream, N, tu.depth, tu.poses, sl, d_ligs, d_meta + s, n, group, ctx.w_A, t.cmax,
                         t.boxmax,
                         prot.spec_lb.as<float2>(), prot.nch, ctx.w_T, 1.0f / (float)G, tw),
              "mid");
    timer.end(e);
    e = timer.begin(3);
    hip_check(launch_inv_argmin(ctx.stream, N, tu.inv, d_meta + s, n, N - 2, ctx.w_T, d_keys + s, tw),
              "inverse + argmin");
    timer.end(e);
  }
  // join: the last two sets' z-inverse stages (earlier ones are ordered
  // before later mid stages) back into the launch stream
  if (ovl) {
    for (size_t si = last_set > 0 ? last_set - 1 : 0; si <= last_set; ++si)
      hip_check(hipStreamWaitEvent(ctx.stream, ctx.set_events[2 * si + 1], 0), "join");
  }
}

}  // namespace

Prepared::~Prepared() {
  drop_graph();
  if (h_up) (void)hipHostFree(h_up);
  if (h_keys) (void)hipHostFree(h_keys);
  if (h_meta) (void)hipHostFree(h_meta);
  if (start) (void)hipEventDestroy(start);
  if (done) (void)hipEventDestroy(done);
}

void Prepared::drop_graph() {
  if (graph) (void)hipGraphExecDestroy(graph);
  graph = nullptr;
}

void build_search(const Protein& prot, const std::vector<Job>& jobs, Prepared& t) {
  t.drop_graph();
  t.ligs.clear();
  t.chan.clear();
  t.pad.clear();
  t.wv.clear();
  t.wq.clear();
  t.coords.clear();
  t.slots.clear();
  t.job_first.clear();
  t.job_n.clear();
  t.astride = 0;
  t.cmax = 0;
  t.boxmax = 2;
  t.max_chan_atoms = 1;
  build_tables(prot, jobs, t);
  for (const Job& j : jobs) t.job_n.push_back(j.npose());
  t.prot = &prot;
}

namespace {

// t's tables to the device on the context stream (context lock held).
void upload_tables(Context& ctx, Prepared& t) {
  if (t.own_staging) {
    // Pipelined callers: the tables go through this search's own pinned
    // staging, so the copies are asynchronous (a copy from pageable memory
    // would hold the host until the stream's earlier work is done).  The
    // staging is rewritten only after finish_search of this search.
    struct Part {
      DevBuf* dst;
      const void* src;
      size_t bytes;
    };
    const Part parts[7] = {{&t.d_ligs, t.ligs.data(), t.ligs.size() * sizeof(LigandDesc)},
                           {&t.d_chan, t.chan.data(), t.chan.size() * sizeof(int)},
                           {&t.d_pad, t.pad.data(), t.pad.size() * sizeof(int)},
                           {&t.d_wv, t.wv.data(), t.wv.size() * sizeof(float)},
                           {&t.d_wq, t.wq.data(), t.wq.size() * sizeof(float)},
                           {&t.d_coords, t.coords.data(), t.coords.size() * sizeof(float)},
                           {&t.d_slots, t.slots.data(), t.slots.size() * sizeof(PoseSlot)}};
    size_t off[7], total = 0;
    for (int i = 0; i < 7; ++i) {
      off[i] = total;
      total += (parts[i].bytes + 255) & ~(size_t)255;
    }
    if (total > t.h_up_cap) {
      if (t.h_up) (void)hipHostFree(t.h_up);
      t.h_up = nullptr;
      t.h_up_cap = 0;
      hip_check(hipHostMalloc(&t.h_up, total), "hipHostMalloc");
      t.h_up_cap = total;
    }
    char* h = static_cast<char*>(t.h_up);
    for (int i = 0; i < 7; ++i) {
      parts[i].dst->ensure(std::max<size_t>(parts[i].bytes, 4));
      if (parts[i].bytes == 0) continue;
      std::memcpy(h + off[i], parts[i].src, parts[i].bytes);
      hip_check(hipMemcpyAsync(parts[i].dst->p, h + off[i], parts[i].bytes, hipMemcpyHostToDevice, ctx.stream),
                "upload");
    }
  } else {
    upload(ctx, t.d_ligs, t.ligs);
    upload(ctx, t.d_chan, t.chan);
    upload(ctx, t.d_pad, t.pad);
    upload(ctx, t.d_wv, t.wv);
    upload(ctx, t.d_wq, t.wq);
    upload(ctx, t.d_coords, t.coords);
    upload(ctx, t.d_slots, t.slots);
  }
  const size_t total = std::max<size_t>(t.slots.size(), 1);
  t.d_keys.ensure(total * sizeof(unsigned long long));
  t.d_meta.ensure(total * sizeof(RotMeta));
}

}  // namespace

void prepare_search(Context& ctx, const Protein& prot, const std::vector<Job>& jobs, Prepared& t) {
  build_search(prot, jobs, t);
  std::lock_guard<std::mutex> lock(ctx.mu);
  hip_check(hipSetDevice(ctx.device), "hipSetDevice");
  upload_tables(ctx, t);
  hip_check(hipStreamSynchronize(ctx.stream), "prepare sync");
}

namespace {

// Enqueues t's launch sequence and the copies of its keys and per-pose
// metadata to host staging (t's own when t.own_staging, else the context's);
// context lock held.
void enqueue_search(Context& ctx, Prepared& t, Engine engine, bool use_graph, EventTimer& timer,
                    hipEvent_t start = nullptr) {
  const Protein& prot = *t.prot;
  const size_t total = t.slots.size();
  if (engine == kEngineReference) {
    if (start) hip_check(hipEventRecord(start, ctx.stream), "hipEventRecord");
    enqueue_reference(ctx, prot, t, timer);
  } else {
    const size_t chunk = fused_chunk(ctx, prot, t);
    ensure_fused_work(ctx, prot, t, chunk);
    (void)ctx.twiddle(prot.N);
    const Tuned tu = tune_fused(ctx, prot, t, chunk);
    ctx.last_fwd = tu.fwd;
    ctx.last_inv = tu.inv;
    ctx.last_pipe = tu.pipe;
    ctx.last_group = tu.mid_group();
    ctx.last_depth = tu.depth;
    ctx.last_poses = tu.poses;
    const int ovl = overlap_z_stage(ctx, tu) ? 1 : 0;
    if (ovl) ensure_overlap(ctx, t, chunk);
    if (start) hip_check(hipEventRecord(start, ctx.stream), "hipEventRecord");
    // Jobs of more launch sets than kMaxGraphSets run eagerly: a host
    // segmentation fault inside the search of a 3.28 M-pose uneven batch
    // captured as one graph of 3 410 launch sets (3 304 had passed) has no
    // cause found in this code, so graphs stay at sizes that have run.
    const size_t nsets = (total + chunk - 1) / chunk;
    if (use_graph && !ctx.time_stages && nsets <= kMaxGraphSets) {
      if (!t.graph || t.graph_gen != ctx.ws_generation || t.graph_chunk != chunk || t.graph_group != tu.mid_group() ||
          t.graph_fwd != tu.fwd || t.graph_inv != tu.inv || t.graph_depth != tu.depth ||
          t.graph_poses != tu.poses || t.graph_pipe != tu.pipe || t.graph_products != ctx.atom_products ||
          t.graph_overlap != ovl) {
        t.drop_graph();
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
        try {
          enqueue_fused(ctx, prot, t, chunk, tu, timer);
        } catch (...) {
          (void)hipStreamEndCapture(ctx.stream, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        hip_check(hipStreamEndCapture(ctx.stream, &g), "hipStreamEndCapture");
        hipError_t ie = hipGraphInstantiate(&t.graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(ie, "hipGraphInstantiate");
        t.graph_gen = ctx.ws_generation;
        t.graph_chunk = chunk;
        t.graph_group = tu.mid_group();
        t.graph_pipe = tu.pipe;
        t.graph_products = ctx.atom_products;
        t.graph_overlap = ovl;
        t.graph_fwd = tu.fwd;
        t.graph_inv = tu.inv;
        t.graph_depth = tu.depth;
        t.graph_poses = tu.poses;
      }
      hip_check(hipGraphLaunch(t.graph, ctx.stream), "hipGraphLaunch");
    } else {
      enqueue_fused(ctx, prot, t, chunk, tu, timer);
    }
  }
  void* hk = ctx.h_keys;
  void* hm = ctx.h_meta;
  if (t.own_staging) {
    if (total > t.h_cap) {
      if (t.h_keys) (void)hipHostFree(t.h_keys);
      if (t.h_meta) (void)hipHostFree(t.h_meta);
      t.h_keys = t.h_meta = nullptr;
      t.h_cap = 0;
      hip_check(hipHostMalloc(&t.h_keys, total * sizeof(unsigned long long)), "hipHostMalloc");
      hip_check(hipHostMalloc(&t.h_meta, total * sizeof(RotMeta)), "hipHostMalloc");
      t.h_cap = total;
    }
    hk = t.h_keys;
    hm = t.h_meta;
  } else {
    ctx.ensure_host(total);
    hk = ctx.h_keys;
    hm = ctx.h_meta;
  }
  hip_check(hipMemcpyAsync(hk, t.d_keys.p, total * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx.stream),
            "keys D2H");
  hip_check(hipMemcpyAsync(hm, t.d_meta.p, total * sizeof(RotMeta), hipMemcpyDeviceToHost, ctx.stream), "meta D2H");
}

// Per-pose results of t from host staging (keys, metadata).
void decode_results(Prepared& t, const void* hkeys, const void* hmeta) {
  const Protein& prot = *t.prot;
  const unsigned long long* keys = static_cast<const unsigned long long*>(hkeys);
  const RotMeta* meta = static_cast<const RotMeta*>(hmeta);
  const int N = prot.N;
  t.results.assign(t.job_first.size(), PoseResults());
  for (size_t j = 0; j < t.job_first.size(); ++j) {
    const size_t first = t.job_first[j];
    const size_t n = t.job_n[j];
    PoseResults& r = t.results[j];
    r.energy.resize(n);
    r.index.resize(n * 3);
    r.mincoor.resize(n * 3);
    r.box.resize(n * 3);
    r.win.resize(n * 3);
    for (size_t i = 0; i < n; ++i) {
      const unsigned long long k = keys[first + i];
      const RotMeta& m = meta[first + i];
      if (k == ~0ull || !m.valid) {
        r.energy[i] = std::numeric_limits<float>::infinity();
        r.index[i * 3 + 0] = r.index[i * 3 + 1] = r.index[i * 3 + 2] = -1;
      } else {
        const unsigned lin = key_index(k);
        r.energy[i] = key_energy(k);
        r.index[i * 3 + 0] = (int)(lin / (unsigned)(N * N));
        r.index[i * 3 + 1] = (int)((lin / (unsigned)N) % (unsigned)N);
        r.index[i * 3 + 2] = (int)(lin % (unsigned)N);
      }
      for (int k3 = 0; k3 < 3; ++k3) {
        r.mincoor[i * 3 + k3] = m.minc[k3];
        r.box[i * 3 + k3] = m.box[k3];
        r.win[i * 3 + k3] = m.win[k3];
      }
    }
  }
}

}  // namespace

void execute_search(Context& ctx, Prepared& t, Engine engine, bool use_graph) {
  std::lock_guard<std::mutex> lock(ctx.mu);
  hip_check(hipSetDevice(ctx.device), "hipSetDevice");
  if (!t.prot) throw Error(kErrInvalid, "search not prepared");
  t.results.assign(t.job_first.size(), PoseResults());
  if (t.slots.empty()) return;
  EventTimer timer(ctx);
  t.own_staging = false;
  enqueue_search(ctx, t, engine, use_graph, timer);
  hip_check(hipStreamSynchronize(ctx.stream), "search sync");
  timer.collect();
  // Decode while the pinned staging is ours (context lock held).
  decode_results(t, ctx.h_keys, ctx.h_meta);
}

void submit_search(Context& ctx, Prepared& t) {
  std::lock_guard<std::mutex> lock(ctx.mu);
  hip_check(hipSetDevice(ctx.device), "hipSetDevice");
  if (!t.prot) throw Error(kErrInvalid, "search not built");
  if (!t.start) hip_check(hipEventCreate(&t.start), "hipEventCreate");
  if (!t.done) hip_check(hipEventCreate(&t.done), "hipEventCreate");
  t.own_staging = true;
  upload_tables(ctx, t);
  if (t.slots.empty()) {
    hip_check(hipEventRecord(t.start, ctx.stream), "hipEventRecord");
  } else {
    EventTimer timer(ctx);  // stages are timed only by the synchronous path
    const bool ts = ctx.time_stages;
    ctx.time_stages = false;
    try {
      enqueue_search(ctx, t, kEngineFused, false, timer, t.start);
    } catch (...) {
      ctx.time_stages = ts;
      throw;
    }
    ctx.time_stages = ts;
  }
  hip_check(hipEventRecord(t.done, ctx.stream), "hipEventRecord");
}

void finish_search(Context& ctx, Prepared& t) {
  (void)ctx;
  if (!t.done) throw Error(kErrInvalid, "search not submitted");
  hip_check(hipEventSynchronize(t.done), "search wait");
  t.results.assign(t.job_first.size(), PoseResults());
  if (!t.slots.empty()) decode_results(t, t.h_keys, t.h_meta);
}

float search_gpu_ms(Prepared& t) {
  float ms = 0.f;
  if (t.start && t.done) hip_check(hipEventElapsedTime(&ms, t.start, t.done), "hipEventElapsedTime");
  return ms;
}

void run_search(Context& ctx, const Protein& prot, const std::vector<Job>& jobs, Engine engine,
                std::vector<PoseResults>& out) {
  Prepared t;
  prepare_search(ctx, prot, jobs, t);
  execute_search(ctx, t, engine, false);
  out = std::move(t.results);
}

namespace {

// Per-atom protein channel (vdW probe of the atom, -1 if none) and weights.
void atom_channels(const Protein& prot, const Ligand& lig, std::vector<int>& pch) {
  pch.assign(lig.natom, -1);
  for (int a = 0; a < lig.natom; ++a) {
    const int c = lig.atom_channel[a];
    if (c >= 0 && c < lig.num_vdw_used) pch[a] = lig.idx_vdw_used[c];
  }
  (void)prot;
}

struct PoseScratch {
  DevBuf coords, pch, wv, wq, energy, grad, inside;
};

void stage_pose_inputs(Context& ctx, const Protein& prot, const Ligand& lig, int npose, const double* coords,
                       PoseScratch& w) {
  std::vector<int> pch;
  atom_channels(prot, lig, pch);
  std::vector<float> wv(lig.natom), wq(lig.natom);
  for (int a = 0; a < lig.natom; ++a) {
    wv[a] = std::sqrt(lig.eps[a]);
    wq[a] = lig.charge[a];
  }
  upload(ctx, w.pch, pch);
  upload(ctx, w.wv, wv);
  upload(ctx, w.wq, wq);
  const size_t n = (size_t)npose * lig.natom * 3;
  w.coords.ensure(std::max<size_t>(n, 1) * sizeof(double));
  if (n) hip_check(hipMemcpyAsync(w.coords.p, coords, n * sizeof(double), hipMemcpyHostToDevice, ctx.stream), "coords");
  w.energy.ensure(std::max(npose, 1) * sizeof(double));
}

}  // namespace

void rescore_poses(Context& ctx, const Protein& prot, const Ligand& lig, int npose, const double* coords,
                   double* energy, double* grad, int* inside) {
  std::lock_guard<std::mutex> lock(ctx.mu);
  hip_check(hipSetDevice(ctx.device), "hipSetDevice");
  if (npose <= 0) return;
  PoseScratch w;
  stage_pose_inputs(ctx, prot, lig, npose, coords, w);
  const size_t n = (size_t)npose * lig.natom * 3;
  if (grad) w.grad.ensure(n * sizeof(double));
  if (inside) w.inside.ensure((size_t)npose * sizeof(int));
  const double gmin[3] = {prot.grid_min(0), prot.grid_min(1), prot.grid_min(2)};
  hip_check(launch_rescore(ctx.stream, npose, lig.natom, w.coords.as<double>(), w.pch.as<int>(), w.wv.as<float>(),
                           w.wq.as<float>(), prot.grid.as<float>(), prot.N, prot.nch - 1, gmin, (double)prot.h,
                           w.energy.as<double>(), grad ? w.grad.as<double>() : nullptr,
                           inside ? w.inside.as<int>() : nullptr),
            "rescore");
  hip_check(hipMemcpyAsync(energy, w.energy.p, (size_t)npose * sizeof(double), hipMemcpyDeviceToHost, ctx.stream),
            "energy D2H");
  if (grad)
    hip_check(hipMemcpyAsync(grad, w.grad.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx.stream), "grad D2H");
  if (inside)
    hip_check(hipMemcpyAsync(inside, w.inside.p, (size_t)npose * sizeof(int), hipMemcpyDeviceToHost, ctx.stream),
              "inside D2H");
  hip_check(hipStreamSynchronize(ctx.stream), "rescore sync");
}

void refine_poses(Context& ctx, const Protein& prot, const Ligand& lig, int npose, double* coords, int steps,
                  double penalty, double* energy) {
  std::lock_guard<std::mutex> lock(ctx.mu);
  hip_check(hipSetDevice(ctx.device), "hipSetDevice");
  if (npose <= 0) return;
  PoseScratch w;
  stage_pose_inputs(ctx, prot, lig, npose, coords, w);
  const size_t n = (size_t)npose * lig.natom * 3;
  const double gmin[3] = {prot.grid_min(0), prot.grid_min(1), prot.grid_min(2)};
  hip_check(launch_refine(ctx.stream, npose, lig.natom, w.coords.as<double>(), w.pch.as<int>(), w.wv.as<float>(),
                          w.wq.as<float>(), prot.grid.as<float>(), prot.N, prot.nch - 1, gmin, (double)prot.h, steps,
                          penalty, w.energy.as<double>()),
            "refine");
  hip_check(hipMemcpyAsync(coords, w.coords.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx.stream), "coords D2H");
  hip_check(hipMemcpyAsync(energy, w.energy.p, (size_t)npose * sizeof(double), hipMemcpyDeviceToHost, ctx.stream),
            "energy D2H");
  hip_check(hipStreamSynchronize(ctx.stream), "refine sync");
}

int generate_conformers(Context& ctx, int natom, const double* coords, int nrotor, const int* rotors,
                        const unsigned char* moving, const unsigned char* exclude, int ntrial, uint64_t seed,
                        double clash_dist, int max_keep, double* out, int* trial_index) {
  std::lock_guard<std::mutex> lock(ctx.mu);
  hip_check(hipSetDevice(ctx.device), "hipSetDevice");
  if (ntrial <= 0 || max_keep <= 0) return 0;
  for (int r = 0; r < nrotor; ++r)
    for (int k = 0; k < 4; ++k)
      if (rotors[r * 4 + k] < 0 || rotors[r * 4 + k] >= natom) throw Error(kErrInvalid, "rotor atom out of range");
  std::vector<double> x(coords, coords + (size_t)natom * 3);
  std::vector<int> rot(rotors, rotors + (size_t)nrotor * 4);
  std::vector<unsigned char> mv(moving, moving + (size_t)nrotor * natom);
  std::vector<unsigned char> ex(exclude, exclude + (size_t)natom * natom);
  DevBuf d_x, d_rot, d_mv, d_ex, d_out, d_ok;
  upload(ctx, d_x, x);
  upload(ctx, d_rot, rot);
  upload(ctx, d_mv, mv);
  upload(ctx, d_ex, ex);
  const size_t per = (size_t)natom * 3;
  d_out.ensure((size_t)ntrial * per * sizeof(double));
  d_ok.ensure((size_t)ntrial * sizeof(int));
  hip_check(launch_conformers(ctx.stream, natom, d_x.as<double>(), nrotor, d_rot.as<int>(),
                              d_mv.as<unsigned char>(), d_ex.as<unsigned char>(), ntrial, seed, clash_dist,
                              d_out.as<double>(), d_ok.as<int>()),
            "conformers");
  std::vector<int> ok(ntrial);
  std::vector<double> all((size_t)ntrial * per);
  hip_check(hipMemcpyAsync(ok.data(), d_ok.p, ok.size() * sizeof(int), hipMemcpyDeviceToHost, ctx.stream), "ok D2H");
  hip_check(hipMemcpyAsync(all.data(), d_out.p, all.size() * sizeof(double), hipMemcpyDeviceToHost, ctx.stream),
            "conformers D2H");
  hip_check(hipStreamSynchronize(ctx.stream), "conformers sync");
  int kept = 0;
  for (int t = 0; t < ntrial && kept < max_keep; ++t) {
    if (!ok[t]) continue;
    std::memcpy(out + (size_t)kept * per, all.data() + (size_t)t * per, per * sizeof(double));
    if (trial_index) trial_index[kept] = t;
    ++kept;
  }
  return kept;
}

}  // namespace cdk
