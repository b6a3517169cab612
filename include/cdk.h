/*
 * cdk.h - C-ABI of the MI355X (gfx950) FFT translational-rotational pose
 * search.  Plain pointers and sizes only; every entry point returns a status
 * code (CDK_OK on success) and never exits the process.  The message of the
 * last failure on the calling thread is available from cdk_last_error().
 *
 * Layout conventions (those of the reference's search drivers):
 *   grids       float [grid][X][Y][Z], Z fastest (ReadGrids.cpp:54-66)
 *   spectra     float [grid][X][Y][Z/2+1][2] (cuFFT R2C layout,
 *               TranRotaConfSearchNoMini.cu:158-174)
 *   coordinates float [atom][3] (Angstrom)
 *   quaternions float [q][4] = (w, x, y, z) (QuaternionUniformSampling.cpp:19-22)
 *   ligand grid float [q][channel][X][Y][Z], channel q*C + c
 *               (FillLigandGrid.cpp:38-39, 92-93)
 *
 * Threading: a context serialises its calls with an internal lock; use one
 * context per host thread for concurrency.  A server object is safe to call
 * from any thread.
 */
#ifndef CDK_H
#define CDK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CDK_OK 0
#define CDK_ERR_INVALID 1     /* bad argument */
#define CDK_ERR_HIP 2         /* HIP runtime failure */
#define CDK_ERR_NOMEM 3       /* device or host allocation failed */
#define CDK_ERR_UNSUPPORTED 4 /* grid edge or size outside the compiled set */
#define CDK_ERR_PENDING 5     /* serving request not finished */
#define CDK_ERR_BUDGET 6      /* request larger than the server's HBM budget */
#define CDK_ERR_CANCELLED 7   /* serving request cancelled */

#define CDK_ENGINE_REFERENCE 0 /* reference dataflow: full-grid stages, batches of B rotations */
#define CDK_ENGINE_FUSED 1     /* pruned, fused passes, protein spectra reused across poses */

typedef struct cdk_context cdk_context;
typedef struct cdk_protein cdk_protein;
typedef struct cdk_ligand cdk_ligand;
typedef struct cdk_batch cdk_batch;
typedef struct cdk_server cdk_server;
typedef struct cdk_rng cdk_rng;

/* ---- library / context ------------------------------------------------ */
const char* cdk_last_error(void);
int cdk_version(void);
/* Device context: one HIP stream, twiddle tables, reusable work arrays.
 * Replaces the cudaMalloc/cufftPlanMany set-up of
 * TranRotaConfSearchNoMini.cu:156-252. */
int cdk_context_create(int device, cdk_context** out);
int cdk_context_destroy(cdk_context* ctx);
/* Options: "ws_budget_bytes", "ref_batch" (B, default 50 as NoMini:208),
 * "fused_chunk" (poses per fused launch set, 0 = from budget), "mid_group"
 * (poses per mid-stage workgroup, 0 = autotuned with the ring depth),
 * "mid_depth" (register buffers in the mid stage's (pose, channel) ring when
 * mid_group is fixed: 2 or 3 at 32^3/64^3, 0 = 2), "mid_poses" (poses of
 * one ligand that share each protein spectrum load when mid_group is fixed:
 * 1, or 2 or 4 at 32^3/64^3; 0 = 1), "pipeline" (fused engine form: 0 plane
 * form, whose forward stage writes ligand planes after the z and y
 * transforms; 1 column form at 64^3, whose forward stage stops after the z
 * transform and whose mid stage runs the y and x transforms per kz plane;
 * 2 atom form at 64^3, 96^3 and 128^3, whose mid stage builds each channel's ligand
 * spectrum per kz plane from the atoms as a sum of outer products on the
 * matrix cores, with no forward stage; -1 = autotuned once per grid size),
 * "atom_products" (the atom form's products: 2 = FP16 matrix cores on
 * operands split exactly into three FP16 parts each (after power-of-two
 * scaling), the six partial products down to 2^-22 with FP32 accumulation:
 * an FP32 x FP32 product to about 2^-32, below FP32 rounding; 0 = two FP16
 * parts each, three partial products, about 2^-21 relative error per
 * product; 1 = FP32 matrix cores; default 2), "overlap_z" (atom form: 1 = each
 * launch set's z-inverse stage runs on a second stream beside the next set's
 * mid stage, with two T2 buffers; 0 = one stream; default 1),
 * "fwd_variant" (forward
 * plane stage: 0 block-wide LDS Stockham FFTs, 1 register four-step FFTs,
 * -1 = autotuned once per grid size), "inv_variant" (inverse
 * plane stage + argmin: 0 block-wide LDS FFTs per plane, 1 register FFTs over
 * 4 planes per workgroup, -1 = autotuned once per grid size), "use_graph",
 * "time_stages". */
int cdk_context_set_option(cdk_context* ctx, const char* key, int64_t value);
/* Reads an option back; also "last_fwd_variant", "last_inv_variant",
 * "last_pipeline", "last_mid_group", "last_mid_depth" and "last_mid_poses"
 * (what the autotuner chose for the last fused search), "max_dyn_lds" (the
 * device's dynamic LDS limit per workgroup, bytes) and "workspace_bytes"
 * (device bytes of the fused engine's launch-set working arrays: one pool,
 * carved per search; never above "ws_budget_bytes" unless "fused_chunk"
 * forces larger launch sets).  No reference counterpart. */
int cdk_context_get_option(cdk_context* ctx, const char* key, int64_t* value);
/* Per-stage device time (HIP events on the context stream) accumulated while
 * "time_stages" is on: ms[8], launches[8].  Stage ids: reference engine
 * 0 prep, 1 spread, 2 R2C, 3 ConjMult, 4 SumGrids, 5 C2R, 6 argmin; fused
 * engine 0 prep, 1 forward planes, 2 mid lines, 3 inverse planes + argmin. */
int cdk_context_stage_times(cdk_context* ctx, double* ms, int* launches, int reset);

/* ---- protein: resident potential grids and spectra ------------------- */
/* Uploads num_grids grids (num_grids-1 vdW probe grids then the elec grid,
 * ReadGrids.cpp:41-66) and computes their forward R2C spectra on the device
 * (TranRotaConfSearchNoMini.cu:177-199).  mid/len as in the grid file header
 * (ReadGrids.cpp:27-38); grid origin = mid - len/2 (NoMini:149-151).
 * Cubic grids with edge 32, 48, 64, 96 or 128. */
int cdk_protein_create(cdk_context* ctx, int num_grids, int xdim, int ydim, int zdim, const float* mid3,
                       const float* len3, float spacing, const float* grid_radii, const float* values,
                       cdk_protein** out);
/* Builds the potential grids on the device from the protein's atoms and
 * keeps them resident like cdk_protein_create (the step before the search:
 * the reference reads grids made elsewhere, ReadGrids.cpp:6-68).  xyz
 * [natom][3] A, charge e, epsilon kcal/mol, radius A.  Grid channel p <
 * num_probes: soft-core LJ sum_i sqrt(eps_i) ((R/d)^12 - 2 (R/d)^6), R = r_i +
 * probe_radii[p], over d <= lj_cutoff, capped at +lj_cap; last channel:
 * Coulomb sum_i 332.0636 q_i / (4 d^2) over d <= elec_cutoff, clamped to
 * +-elec_cap; d^2 floored at 1e-4.  Each grid point sums its atoms in
 * ascending index order in FP64.  Grid point (i,j,k) at mid - len/2 +
 * (i,j,k) spacing; cubic, edge n as for cdk_protein_create. */
int cdk_protein_build(cdk_context* ctx, int natom, const double* xyz, const double* charge, const double* epsilon,
                      const double* radius, int num_probes, const float* probe_radii, int n, float spacing,
                      const float* mid3, const float* len3, double lj_cutoff, double elec_cutoff, double lj_cap,
                      double elec_cap, cdk_protein** out);
/* Copies the resident grids out: float [num_grids][X][Y][Z]. */
int cdk_protein_grids(cdk_protein* p, float* values);
int cdk_protein_destroy(cdk_protein* p);
/* Copies the spectrum of one grid channel out: float [X][Y][Z/2+1][2]. */
int cdk_protein_spectrum(cdk_protein* p, int channel, float* out);
/* Device memory the protein holds (grids + spectra), bytes. */
int cdk_protein_bytes(cdk_protein* p, int64_t* bytes);

/* ---- ligand ----------------------------------------------------------- */
/* Ligand force-field parameters (charge, epsilon in kcal/mol, radius in A,
 * as GetNonbondedParameters.cpp:5-35 produces) with the channel assignment
 * of GetIdxOfAtomsForVdwRadius.cpp:4-35 against the protein's probe radii. */
int cdk_ligand_create(cdk_protein* p, int natom, const float* charge, const float* epsilon, const float* radius,
                      cdk_ligand** out);
int cdk_ligand_destroy(cdk_ligand* l);
/* num_vdw_used, idx_vdw_used[num_grids-1] (first num_vdw_used valid) and the
 * probe-radius index of each atom (GetIdxOfAtomsForVdwRadius.cpp:13-21). */
int cdk_ligand_channels(const cdk_ligand* l, int* num_vdw_used, int* idx_vdw_used, int* atom_radius_index);

/* ---- search: the TranRotaConfSearch hot loop ------------------------- */
/* For every conformer c and quaternion r: rotate, spread, correlate with the
 * resident protein spectra over all translations and keep the windowed
 * minimum (TranRotaConfSearchNoMini.cu:297-464).  quats is
 * [nconf][nquat][4] (redrawn per conformer, NoMini:321-324) or, when
 * shared_quats != 0, one [nquat][4] set for all conformers.
 * Outputs (caller-owned, [nconf*nquat]): energy (+inf when the window is
 * empty), index [.][3] best translation (i, j, k) (-1 when empty),
 * mincoor [.][3] minimum corner of the rotated ligand (GetMinCoors.cpp). */
int cdk_search(cdk_context* ctx, cdk_protein* p, const cdk_ligand* l, int nconf, const float* conf_coords,
               int nquat, const float* quats, int shared_quats, int engine, float* energy, int* index,
               float* mincoor);

/* Ranks rotations by energy and rebuilds the lowest n poses of one conformer
 * (sort_index, GeneDiverseConformations.h:9-20, ties by rotation index;
 * pose = rotated - mincoor + gridMin + index*spacing, NoMini:466-487).
 * rotated: [nquat][natom][3] coordinates of every rotation (from
 * cdk_rotate); out_rank[n] rotation indices, out_pose [n][natom][3] double. */
int cdk_rank_poses(cdk_protein* p, int natom, int nquat, const float* rotated, const float* energy,
                   const int* index, const float* mincoor, int n_lowest, int* out_rank, double* out_pose);

/* cdk_rank_poses from the conformer and the rotations themselves (what the
 * driver holds at NoMini:476-480) instead of pre-rotated coordinates:
 * conformer [natom][3], quats [nquat][4]; only the kept rotations are
 * rotated (Rotate.cpp:4-27, bit-identical to the search's rotation). */
int cdk_rank_poses_quats(cdk_protein* p, int natom, int nquat, const float* conformer, const float* quats,
                         const float* energy, const int* index, const float* mincoor, int n_lowest, int* out_rank,
                         double* out_pose);

/* ---- retained poses: grid rescoring and rigid refinement ------------- */
/* Grid score of explicit poses, in FP64: coords [npose][natom][3] absolute
 * (A, the frame of the poses cdk_rank_poses rebuilds); energy [npose] =
 * sum_a sqrt(eps_a) P_probe(a)(r_a) + q_a P_elec(r_a), P trilinear in the
 * protein grids (the interpolation that the search's spreading implies, so a
 * pose from the search scores its search energy).  grad [npose][natom][3]
 * (dE/dr, kcal/mol/A) and n_inside [npose] (atoms on the grid; off-grid atoms
 * add nothing) may be NULL.  Replaces the grid energy of the OpenMM grid
 * context (TranRotaConfSearchGrid.cu:158-174, AddGridForcesToOpenMMSystem.cpp
 * :8-130), which interpolates with splines instead. */
int cdk_rescore_poses(cdk_protein* p, const cdk_ligand* l, int npose, const double* coords, double* energy,
                      double* grad, int* n_inside);
/* Rigid-body descent on that score, in place on coords: per pose `steps`
 * trial moves along the force and the torque about the centroid, kept when
 * the score drops.  An atom off the grid adds off_grid_penalty.  energy
 * [npose] receives the final score.  A fast pre-minimisation before the host
 * OpenMM step (Grid:511-532). */
int cdk_refine_poses(cdk_protein* p, const cdk_ligand* l, int npose, double* coords, int steps,
                     double off_grid_penalty, double* energy);

/* Writes n retained poses of one conformer as conformer_<conformer>_<i>.pdb
 * in dir (HETATM records; atom_names / elements may be NULL) and appends one
 * row "conformer_<c>_<i>.pdb,<c>,<i>,<energy>" per pose to energy_path (if
 * not NULL), the drivers' output (TranRotaConfSearchNoMini.cu:489-502).
 * Host only. */
int cdk_write_poses(const char* dir, int conformer, int n, int natom, const double* poses, const float* energy,
                    const char* const* atom_names, const char* const* elements, const char* energy_path);
/* One protein grid channel in the order OpenMM's Continuous3DFunction takes
 * (x fastest, double; AddGridForcesToOpenMMSystem.cpp:30-41) and the grid
 * bounds in nm [min x, y, z, max x, y, z] (TranRotaConfSearchGrid.cu:149-154).
 * values (X*Y*Z doubles) or bounds_nm may be NULL. */
int cdk_grid_for_openmm(cdk_protein* p, int channel, double* values, double* bounds_nm);

/* ---- conformers -------------------------------------------------------- */
/* Random-torsion conformers on the device (GeneRandomConformations.cpp:36-47;
 * the OpenMM minimisation and energy cap of :49-109 are replaced by a steric
 * filter).  Trial t < ntrial sets each rotor r = (a, b, c, d), in order, to
 * the IUPAC dihedral theta = -pi + 2 pi u(seed, t, r) by turning the atoms
 * flagged in moving[r][natom] about b->c; a trial is kept when no atom pair
 * with exclude[i][j] == 0 is closer than clash_dist.  The first max_keep
 * kept trials, in trial order, go to out [max_keep][natom][3] (and their
 * trial numbers to trial_index, may be NULL); *n_kept receives the count. */
int cdk_generate_conformers(cdk_context* ctx, int natom, const double* coords, int nrotor, const int* rotors,
                            const uint8_t* moving, const uint8_t* exclude, int ntrial, uint64_t seed,
                            double clash_dist, int max_keep, double* out, int* trial_index, int* n_kept);

/* ---- batched search: many ligands in one set of launches ------------- */
int cdk_batch_create(cdk_context* ctx, cdk_protein* p, cdk_batch** out);
int cdk_batch_add(cdk_batch* b, const cdk_ligand* l, int nconf, const float* conf_coords, int nquat,
                  const float* quats, int shared_quats, int* job_id);
int cdk_batch_run(cdk_batch* b, int engine);
int cdk_batch_poses(cdk_batch* b, int64_t* total_poses);
int cdk_batch_results(cdk_batch* b, int job_id, float* energy, int* index, float* mincoor);
/* Per pose of a job: occupied ligand box [.][3] (max grid index + 2) and
 * translation window [.][3]; the sizes the pruned passes work on. */
int cdk_batch_geometry(cdk_batch* b, int job_id, int* box, int* windows);
int cdk_batch_destroy(cdk_batch* b);

/* ---- serving: continuous batching over resident proteins ------------ */
/* hbm_budget_bytes covers the context's workspace (its "ws_budget_bytes",
 * reserved at creation; CDK_ERR_BUDGET if the budget does not exceed it) plus
 * the resident state of admitted requests (pose slots, conformer coordinates,
 * quaternions and per-pose results, about 104 B per pose).  Requests are
 * admitted in arrival order while that fits, and their poses are batched
 * with whatever else is in flight (at most max_batch_poses per launch set).
 * Launch sets go least-attained-service first: the admitted request with the
 * fewest GPU seconds served so far picks the protein, and every admitted
 * request of that protein shares the set.  Two launch sets are in flight at a
 * time (the next one built and queued while the current one runs).  Results
 * of a request do not depend on what it was batched with.  One worker thread
 * per server; submit/result/stats may be called from any thread. */
int cdk_server_create(cdk_context* ctx, int nprot, cdk_protein* const* prots, int64_t hbm_budget_bytes,
                      int max_batch_poses, cdk_server** out);
int cdk_server_submit(cdk_server* s, int protein, int natom, const float* charge, const float* epsilon,
                      const float* radius, int nconf, const float* conf_coords, int nquat, const float* quats,
                      int shared_quats, int64_t* request_id);
/* wait_ms < 0 blocks until done; 0 polls.  Returns CDK_ERR_PENDING if not
 * finished.  latency_s: submit -> results available on the host.  The call
 * that returns a finished request's results (or its error) releases the
 * request: its id is unknown afterwards (CDK_ERR_INVALID); of several threads
 * collecting one id, one gets the results and the others CDK_ERR_INVALID. */
int cdk_server_result(cdk_server* s, int64_t request_id, int wait_ms, float* energy, int* index, float* mincoor,
                      double* latency_s);
/* The first n of: requests done, poses done, launch sets, mean poses per
 * launch set, GPU seconds of the launch sets, admitted bytes now, queued
 * requests, peak admitted bytes, requests held (not yet collected), launch
 * sets in flight.  Entries past the tenth are not written. */
int cdk_server_stats(cdk_server* s, int n, double* stats);
int cdk_server_destroy(cdk_server* s);

/* ---- reference pose-preparation surface, on the device --------------- */
/* Rotate.cpp:4-27 for every (quaternion, atom): out [nq][natom][3]. */
int cdk_rotate(cdk_context* ctx, int nq, int natom, const float* quats, const float* coords, float* out);
/* Rotation + GetMinCoors.cpp:4-29 + GetMaxCoors.cpp:3-27 + extents
 * (NoMini:352-372) + translation window for an N^3 grid of spacing h
 * (NoMini:446-450).  Any output pointer may be NULL. */
int cdk_rotate_aabb(cdk_context* ctx, int nq, int natom, const float* quats, const float* coords, int grid_edge,
                    float spacing, float* rotated, float* mincoors, float* lengths, int* windows);
/* FillLigandGrid.cpp:7-120 on the device.  Adds into ligand_grid_values
 * (the caller zeroes it first, as NoMini:389).  The vdW atom lists are given
 * in CSR form: atoms of probe radius r are radius_atoms[radius_offsets[r] ..
 * radius_offsets[r+1]) (idxOfAtomVdwRadius); idx_vdw_used[num_vdw_used]. */
int cdk_fill_ligand_grid(cdk_context* ctx, int nq, int natom, const float* coors, const float* mincoors,
                         const float* charge, const float* epsilon, int num_vdw_used, const int* idx_vdw_used,
                         int num_vdw_grids, const int* radius_offsets, const int* radius_atoms, int xdim, int ydim,
                         int zdim, float spacing, float* ligand_grid_values);
/* Batched unnormalised 3-D R2C / C2R (cuFFT plans of NoMini:185-188,
 * 222-225, 244-247).  c2r's input is not modified. */
int cdk_fft_r2c_3d(cdk_context* ctx, int n, int batch, const float* in, float* out);
int cdk_fft_c2r_3d(cdk_context* ctx, int n, int batch, const float* in, float* out);
/* kernel.cu:5-14 ConjMult: lig[q][c] <- conj(lig[q][c]) * prot[c], in place,
 * H = odist complex values per grid. */
int cdk_conj_mult(cdk_context* ctx, int nq, int nch, int64_t H, const float* prot, float* lig);
/* kernel.cu:16-30 SumGrids: sum[q] = sum_c lig[q][c] / sqrtf(G). */
int cdk_sum_grids(cdk_context* ctx, int nq, int nch, int64_t H, int64_t G, const float* lig, float* sum);
/* NoMini:441-464 windowed argmin over energy grids [nq][n][n][n] with
 * E = f / sqrt(G); windows [nq][3]; out energy[nq], index[nq][3]. */
int cdk_argmin_window(cdk_context* ctx, int nq, int n, const float* grids, const int* windows, float* energy,
                      int* index);

/* No reference counterpart: cdk_argmin_window for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Every pointer is device memory: grids
 * [nq][n][n][n], windows [nq][3] in, energy [nq] and index [nq][3] out.  The
 * per-pose window metadata and argmin keys live in a fixed per-device array
 * (72 B per pose, at most 16384 poses per call: CDK_ERR_INVALID beyond, with
 * nothing launched); the call allocates nothing.  stream: a hipStream_t to
 * enqueue on (the call returns without waiting), or NULL for the context
 * stream (the call returns when the outputs are written).  Calls that overlap
 * in time on one device share that array and race. */
int cdk_argmin_window_noalloc(cdk_context* ctx, int nq, int n, const float* grids, const int* windows,
                              float* energy, int* index, void* stream);

/* ---- host-side input producers --------------------------------------- */
/* std::mt19937_64 generator, as the reference's drivers use (NoMini:203-204,
 * here with an explicit seed). */
int cdk_rng_create(uint64_t seed, cdk_rng** out);
int cdk_rng_destroy(cdk_rng* r);
/* QuaternionUniformSampling.cpp:5-24 (Shoemake): out [n][4]. */
int cdk_quaternion_uniform_sampling(cdk_rng* r, int n, float* out);
/* The same draws from a caller-owned generator: mt19937_64 points to a
 * std::mt19937_64 of a C++ caller built against the same C++ standard
 * library (QuaternionUniformSampling.h:10 takes std::mt19937_64&). */
int cdk_quaternion_uniform_sampling_mt(void* mt19937_64, int n, float* out);
/* Rotate.cpp:4-27 on the host (no context), bit-identical to cdk_rotate:
 * out [nq][natom][3]. */
int cdk_rotate_host(int nq, int natom, const float* quats, const float* coords, float* out);
/* GetMinCoors.cpp:4-29 and GetMaxCoors.cpp:3-27 over rotated coordinates
 * [nq][natom][3] (host); either output may be NULL. */
int cdk_min_max_coors(int nq, int natom, const float* coors, float* mincoors, float* maxcoors);
/* FilterQuaternions.cpp:11-83: keeps quaternions whose rotated extent is
 * below the grid length on every axis; if more than max_used remain, a
 * seeded shuffle (the reference uses rand()-driven random_shuffle) picks
 * them.  Returns the count in *n_used and the kept quaternions in out. */
int cdk_filter_quaternions(cdk_context* ctx, const float* coor, int natom, int nq, const float* quats,
                           float len_x, float len_y, float len_z, int max_used, uint64_t shuffle_seed,
                           int* n_used, float* out);
/* GetIdxOfAtomsForVdwRadius.cpp:4-35: per-atom radius index, the used radii
 * (ascending) and their count. */
int cdk_idx_of_atoms_for_vdw_radius(int natom, const float* atom_radii, int num_vdw_grids, const float* grid_radii,
                                    int* num_vdw_used, int* idx_vdw_used, int* atom_radius_index);

/* ---- seeded synthetic workloads -------------------------------------- */
/* Pseudo-protein: natom atoms uniform in [0, box)^3, charges N(0, 0.3),
 * radii uniform over the 8 probe radii, epsilon uniform [0.05, 0.25];
 * 8 soft-core LJ grids (one per probe radius, capped at +3 kcal/mol) and one
 * Coulomb grid (dielectric 4r, capped at +-20) on an n^3 grid of spacing h
 * centred in the box.  values [9][n][n][n]; radii[8]; mid3, len3. */
/* The atoms behind cdk_synth_protein (same seed, same draws): xyz [natom][3]
 * uniform in [0, box)^3, charge N(0, 0.3), epsilon U[0.05, 0.25], radius
 * from the probe set. */
int cdk_synth_protein_atoms(uint64_t seed, int natom, float box, double* xyz, double* charge, double* epsilon,
                            double* radius);
int cdk_synth_protein(uint64_t seed, int natom, float box, int n, float spacing, float* values, float* radii,
                      float* mid3, float* len3);
/* Random-walk ligand (1.5 A steps, >= 1.2 A between atoms), charges
 * N(0, 0.25) shifted to sum 0, radii from the probe set, epsilon uniform
 * [0.05, 0.20], centred at the origin. */
int cdk_synth_ligand(uint64_t seed, int natom, float* coords, float* charge, float* epsilon, float* radius);
/* Conformers by random torsion perturbation of the walk (conformer 0 is the
 * input), each centred at the origin: out [nconf][natom][3]. */
int cdk_synth_conformers(uint64_t seed, int natom, const float* coords, int nconf, float* out);

/* ---- grid file (ReadGrids.cpp:6-68 text format) ---------------------- */
int cdk_read_grid_header(const char* path, int* num_grids, int* dims3, float* mid3, float* len3, float* spacing,
                         float* restraint_k);
int cdk_read_grids(const char* path, float* radii, float* values);
int cdk_write_grids(const char* path, int num_grids, const int* dims3, const float* mid3, const float* len3,
                    float spacing, float restraint_k, const float* radii, const float* values);

#ifdef __cplusplus
}
#endif

#endif /* CDK_H */
