// Host engine: device context, resident protein spectra, ligand tables and
// the two search pipelines (reference dataflow and fused/pruned).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "cdk_types.hpp"

namespace cdk {

// Error carrying a C-ABI status code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status {
  kOk = 0,
  kErrInvalid = 1,
  kErrHip = 2,
  kErrNoMem = 3,
  kErrUnsupported = 4,
  kErrPending = 5,
  kErrBudget = 6,
  kErrCancelled = 7,
};

void hip_check(hipError_t e, const char* what);

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  void ensure(size_t bytes);
  void release();
  template <class T> T* as() const { return static_cast<T*>(p); }
};

struct Context;

// Resident protein: potential grids and their forward spectra, all channels,
// FP32, [channel][X][Y][Z] and [channel][X][Y][Z/2+1] (NoMini:177-199).
struct Protein {
  Context* ctx = nullptr;
  int nch = 0;  // num grids: vdW channels + 1 elec (ReadGrids.cpp:27)
  int N = 0;
  float h = 0.f;
  float mid[3] = {0, 0, 0};
  float len[3] = {0, 0, 0};
  std::vector<float> radii;  // nch - 1 probe radii
  DevBuf grid;               // float [nch][G]
  DevBuf spec;               // float2 [nch][H]
  DevBuf spec_lb;            // the same spectra in the fused engine's line-blocked layout
  DevBuf spec_kz;            // and in the column form's kz-plane layout (grid edge 64 only)
  DevBuf spec_mf;            // and in the atom form's matrix-core layout (grid edge 64 only)
  size_t G() const { return (size_t)N * N * N; }
  size_t H() const { return (size_t)N * N * (N / 2 + 1); }
  // gridMin = mid - len / 2 evaluated in float, then widened (NoMini:149-151).
  double grid_min(int k) const { return (double)(mid[k] - len[k] / 2); }
};

// Ligand with its channel assignment against a protein's probe radii
// (GetIdxOfAtomsForVdwRadius.cpp:4-35).
struct Ligand {
  int natom = 0;
  std::vector<float> charge, eps, radius;
  int num_vdw_used = 0;
  std::vector<int> idx_vdw_used;               // used probe indices, ascending
  std::vector<std::vector<int>> atoms_of_radius;  // per probe radius, ascending atom indices
  std::vector<int> atom_channel;               // per atom: ligand vdW channel index
  // kernel-side tables for this ligand alone (list_off = w_off = 0)
  LigandDesc desc{};
  std::vector<int> chan_atoms;
  std::vector<int> pad_atoms;  // chan_atoms with each channel padded to a multiple of 8 by -1
  std::vector<float> wv, wq;
  int nch() const { return num_vdw_used + 1; }
};

std::unique_ptr<Ligand> make_ligand(const Protein& p, int natom, const float* charge, const float* eps,
                                    const float* radius);

// One docking job: a ligand, its conformers and per-conformer quaternions.
struct Job {
  const Ligand* lig = nullptr;
  int nconf = 0;
  const float* coords = nullptr;  // [nconf][natom][3]
  int nquat = 0;
  const float* quats = nullptr;   // [nconf][nquat][4]
  bool shared_quats = false;      // one quaternion set [nquat][4] for every conformer
  // Pose range [pose_begin, pose_end) in [conformer][quaternion] order.
  size_t pose_begin = 0;
  size_t pose_end = ~(size_t)0;
  size_t npose() const {
    const size_t n = (size_t)nconf * nquat;
    const size_t e = pose_end < n ? pose_end : n;
    return e > pose_begin ? e - pose_begin : 0;
  }
};

// Per-pose result over a job's pose range, [conformer][quaternion] order.
struct PoseResults {
  std::vector<float> energy;   // +inf when the window is empty
  std::vector<int> index;      // [n][3] best translation (i, j, k); -1 when empty
  std::vector<float> mincoor;  // [n][3]
  std::vector<int> box;        // [n][3] occupied ligand box (max grid index + 2)
  std::vector<int> win;        // [n][3] translation window
};

enum Engine { kEngineReference = 0, kEngineFused = 1 };

struct StageTimes {
  double ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int launches[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// Search tables resident on the device (ligand table, channel lists,
// weights, conformer coordinates, pose slots, results) plus an optional
// hipGraph of the whole fused launch sequence.
struct Prepared {
  const Protein* prot = nullptr;
  std::vector<LigandDesc> ligs;
  std::vector<int> chan, pad;
  std::vector<float> wv, wq, coords;
  std::vector<PoseSlot> slots;
  std::vector<size_t> job_first, job_n;
  int astride = 0, cmax = 0, boxmax = 2, max_chan_atoms = 1;
  int dummy_lig = 0;
  DevBuf d_ligs, d_chan, d_pad, d_wv, d_wq, d_coords, d_slots, d_keys, d_meta;
  hipGraphExec_t graph = nullptr;
  uint64_t graph_gen = 0;
  size_t graph_chunk = 0;
  int graph_group = 0, graph_fwd = -1, graph_inv = -1, graph_depth = 0, graph_poses = 0, graph_pipe = -1,
      graph_products = -1, graph_overlap = -1;
  std::vector<PoseResults> results;
  // Pipelined callers (the server): the tables are uploaded from, and the
  // results copied into, this table's own pinned staging, and `done` marks
  // the end of its launch sequence on the context stream (`start` the first
  // launch), so the host can build the next table while this one runs.
  bool own_staging = false;
  void* h_up = nullptr;  // pinned staging of the tables' upload
  size_t h_up_cap = 0;
  void* h_keys = nullptr;
  void* h_meta = nullptr;
  size_t h_cap = 0;
  hipEvent_t start = nullptr, done = nullptr;
  Prepared() = default;
  Prepared(const Prepared&) = delete;
  Prepared& operator=(const Prepared&) = delete;
  ~Prepared();
  void drop_graph();
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  // The atom form's z-inverse stage of launch set i runs here, overlapped
  // with set i + 1's mid stage on `stream` (two T2 buffers; the ordering
  // events, two per launch set, are created before any capture).
  hipStream_t stream_z = nullptr;
  std::vector<hipEvent_t> set_events;
  std::mutex mu;
  std::map<int, DevBuf*> twiddles;
  // options
  size_t ws_budget = 8ull << 30;  // bytes of per-batch working arrays
  int ref_batch = 50;             // poses per reference-dataflow batch (NoMini:208)
  int fused_chunk = 0;            // 0 = from budget
  int mid_group = 0;              // poses per mid-stage workgroup; 0 = tuned
  int mid_depth = 0;              // mid-stage register ring depth with a fixed group; 0 = 2
  int mid_poses = 0;              // poses sharing each protein load in the mid stage with a fixed group; 0 = 1
  int pipeline = -1;              // fused pipeline form (kPipePlanes, kPipeColumns, kPipeAtoms); -1 = tuned
  int overlap_z = 1;              // atom form: overlap each launch set's z-inverse stage with the next set
  int atom_products = 2;          // atom form products: kAtomF16x3 (exact 3-part FP16 splits), kAtomF16, kAtomF32
  int fwd_variant = -1;           // forward plane stage form; -1 = tuned
  int inv_variant = -1;           // inverse plane stage form; -1 = tuned
  bool use_graph = true;
  bool time_stages = false;
  // Autotuner state.  The forward and inverse plane forms change FFT
  // rounding, so each is chosen once per grid size and kept for the
  // context's life (results then do not depend on how poses are batched);
  // the mid group size only changes the schedule and is tuned per launch
  // shape, with the ring depth and the poses per protein load.
  size_t max_dyn_lds = 64 * 1024;
  std::map<int, int> tuned_pipe;                     // N -> pipeline form
  std::map<int, int> tuned_fwd;                      // N -> variant
  std::map<int, int> tuned_inv;                      // N -> variant
  // (N, cmax, boxmax, chunk rounded up to a power of two) -> (group, depth,
  // poses per protein load)
  std::map<std::array<int, 4>, std::array<int, 3>> tuned_group;
  std::map<std::array<int, 4>, int> tuned_col_group;   // (N, cmax, 0, 0) -> column form group
  std::map<std::array<int, 4>, int> tuned_atom_group;  // (N, cmax, 0, 0) -> atom form group
  int last_fwd = -1, last_inv = -1, last_group = 0, last_depth = 0, last_poses = 0,
      last_pipe = -1;
  StageTimes stage_times;
  // Launch-set working arrays: one pool, never above ws_budget (unless
  // fused_chunk is forced), carved per search into the rotated atoms, atom
  // placements, A (ligand or product spectra), T (correlation spectra, the
  // atom form's T2) and the overlapped atom form's second T2 buffer.
  // ws_generation counts reallocations and re-carvings, so a captured graph
  // that baked in the old addresses is rebuilt.
  DevBuf d_pool;
  float* w_rot = nullptr;
  AtomGrid* w_ag = nullptr;
  float2* w_A = nullptr;
  float2* w_T = nullptr;
  float2* w_T2b = nullptr;
  // the reference dataflow's batch arrays
  DevBuf d_L, d_Lhat, d_S, d_E;
  uint64_t ws_generation = 1;
  // pinned host staging
  void* h_keys = nullptr;
  void* h_meta = nullptr;
  size_t h_cap = 0;

  explicit Context(int dev);
  ~Context();
  const float2* twiddle(int N);
  void ensure_host(size_t nslots);
};

void protein_upload(Protein& p, int nch, int N, float h, const float* mid, const float* len, const float* radii,
                    const float* values);
// Builds the potential grids on the device from atoms (gridbuild.hip), then
// the spectra as protein_upload does.
void protein_build(Protein& p, int natom, const double* xyz, const double* charge, const double* epsilon,
                   const double* radius, int nprobe, const float* probe_radii, int N, float h, const float* mid,
                   const float* len, double lj_cutoff, double elec_cutoff, double lj_cap, double elec_cap);

// Largest job (in launch sets) captured as one hipGraph; larger jobs launch
// eagerly.
constexpr size_t kMaxGraphSets = 2048;
// Plane-form schedules remembered per context; beyond this many the map drops
// its smallest key before adding one.
constexpr size_t kMaxTunedShapes = 256;

// Builds and uploads the tables of a set of jobs (inputs resident in HBM).
void prepare_search(Context& ctx, const Protein& prot, const std::vector<Job>& jobs, Prepared& t);
// Runs every prepared pose; results land in t.results (one entry per job).
// With use_graph the fused launch sequence is captured once and replayed.
void execute_search(Context& ctx, Prepared& t, Engine engine, bool use_graph);
// The same in pipelined pieces: build_search fills t's host tables (no
// device work, no lock); submit_search uploads them and enqueues the fused
// launch sequence and the result copies on the context stream without
// waiting (t.own_staging is set; t's host tables must stay unchanged until
// finish_search); finish_search waits for t's sequence and decodes
// t.results.  Successive submissions run in stream order.
void build_search(const Protein& prot, const std::vector<Job>& jobs, Prepared& t);
void submit_search(Context& ctx, Prepared& t);
void finish_search(Context& ctx, Prepared& t);
// GPU milliseconds of t's last submitted sequence (after finish_search).
float search_gpu_ms(Prepared& t);
// prepare + execute without a graph.
void run_search(Context& ctx, const Protein& prot, const std::vector<Job>& jobs, Engine engine,
                std::vector<PoseResults>& out);

// Grid score of explicit poses (FP64 trilinear, see rescore.hip): coords
// [npose][natom][3] in A; energy [npose]; grad [npose][natom][3] and inside
// [npose] optional.  refine_poses moves each pose rigidly downhill in place.
void rescore_poses(Context& ctx, const Protein& prot, const Ligand& lig, int npose, const double* coords,
                   double* energy, double* grad, int* inside);
void refine_poses(Context& ctx, const Protein& prot, const Ligand& lig, int npose, double* coords, int steps,
                  double penalty, double* energy);

// Random-torsion conformers on the device (conformers.hip): every trial sets
// the rotors in order to angles from a (seed, trial, rotor) stream; trials
// passing the steric filter are kept in trial order, at most max_keep.
// Returns the number kept; out [max_keep][natom][3], trial_index [max_keep].
int generate_conformers(Context& ctx, int natom, const double* coords, int nrotor, const int* rotors,
                        const unsigned char* moving, const unsigned char* exclude, int ntrial, uint64_t seed,
                        double clash_dist, int max_keep, double* out, int* trial_index);

// Retained-pose export (export.cpp).
std::string pose_file_name(int conformer, int rank);
std::string energy_row(int conformer, int rank, float energy);
void write_poses(const char* dir, int conformer, int n, int natom, const double* poses, const float* energy,
                 const char* const* names, const char* const* elements, const char* energy_path);
// One protein channel in OpenMM's x-fastest order (double) and the grid
// bounds in nm: min x, y, z then max x, y, z.
void grid_for_openmm(Context& ctx, const Protein& prot, int channel, double* values, double* bounds_nm);

// Decode helpers shared with tests.
float key_energy(unsigned long long key);
unsigned key_index(unsigned long long key);

}  // namespace cdk
