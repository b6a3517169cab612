// Host-side launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "cdk_types.hpp"

namespace cdk {

bool grid_edge_supported(int N);
// Line groups (four wavefronts of transform lines each) of an N-point
// half spectrum, the plane form's mid-stage work items per pose group.
int mid_line_groups(int N);
// Copies [nch][N][N*(N/2+1)] spectra into the fused engine's line-blocked
// layout (blocks of 64/L lines, rows contiguous within a block).
hipError_t launch_to_line_blocks(hipStream_t st, int N, const float2* in, float2* out, int nch);

hipError_t launch_prep(hipStream_t st, const PoseSlot* slots, int nslots, const LigandDesc* ligs,
                       const float* coords, int N, float h, int astride, int boxmax, float* rot, AtomGrid* ag,
                       RotMeta* meta, unsigned long long* keys);

// Reference-dataflow pipeline stages.
hipError_t launch_spread_full(hipStream_t st, const PoseSlot* slots, int nslots, const LigandDesc* ligs,
                              const int* chan_atoms, const float* wv, const float* wq, const AtomGrid* ag,
                              int astride, const RotMeta* meta, int N, int C, float* out);
hipError_t launch_r2c_3d(hipStream_t st, int N, const float* in, float2* out, int nb, const float2* tw);
hipError_t launch_c2r_3d(hipStream_t st, int N, float2* in, float* out, int nb, const float2* tw);
hipError_t launch_conj_mult(hipStream_t st, float2* L, const float2* P, const int* pmap, int C, long long H,
                            int nb);
hipError_t launch_sum_grids(hipStream_t st, const float2* L, float2* S, int C, long long H, int nb, float G);
hipError_t launch_argmin_full(hipStream_t st, const float* E, const RotMeta* meta, unsigned long long* keys,
                              int N, int nb);

// Fused, pruned pipeline stages.
// Forward plane stage variants: one workgroup per (pose, channel) plane by
// plane, with block-wide LDS Stockham FFTs or with register four-step FFTs.
// The register form needs fwd_reg_lds_bytes() of dynamic LDS;
// configure_fwd_reg() records the device's per-workgroup limit, and
// fwd_reg_supported(N) says whether every box and atom count of grid size N
// fits under it (the variant is then usable for any batch).
constexpr int kFwdBlock = 0, kFwdReg = 1;
size_t fwd_reg_lds_bytes(int N, int boxmax, int max_chan_atoms);
hipError_t configure_fwd_reg(size_t max_dyn_lds);
bool fwd_reg_supported(int N);
hipError_t launch_fwd_planes(hipStream_t st, int N, int variant, const PoseSlot* slots, int nslots,
                             const LigandDesc* ligs, const int* chan_atoms, const float* wv, const float* wq,
                             const AtomGrid* ag, int astride, const RotMeta* meta, int cmax, int boxmax,
                             int max_chan_atoms, float2* A, const float2* tw);
// Mid stage.  np = 1: one pose per wavefront at a time, depth = register
// buffers in the (pose, channel) ring, 2..mid_max_depth(N) (lines of > 8
// values per lane use one buffer).  np = 2 or 4 (<= mid_max_poses(N)): runs
// of np same-ligand poses share each protein load (depth unused).  Every
// (depth, np, group) gives bit-identical results.
int mid_max_depth(int N);
int mid_max_poses(int N);
hipError_t launch_mid(hipStream_t st, int N, int depth, int np, const PoseSlot* slots, const LigandDesc* ligs,
                      const RotMeta* meta, int nslots, int group, const float2* A, int cmax, int boxmax,
                      const float2* Phat, int npch, float2* T, float invG, const float2* tw);
// Column form of the fused pipeline (grid edge 64): the forward stage
// stops after the z transform (Z, kz-major rows of the ligand box), the mid
// stage runs the y and x transforms per kz plane through LDS and writes the
// window rows (T2), the last stage runs the z C2R and the argmin.  P2 is the
// protein spectra in the mid stage's register layout (launch_to_kz_blocks).
constexpr int kPipePlanes = 0, kPipeColumns = 1;
bool col_supported(int N);
hipError_t launch_fwd_z(hipStream_t st, int N, const PoseSlot* slots, int nslots, const LigandDesc* ligs,
                        const int* chan_atoms, const float* wv, const float* wq, const AtomGrid* ag, int astride,
                        const RotMeta* meta, int cmax, int boxmax, int max_chan_atoms, float2* Z, const float2* tw);
hipError_t launch_mid_kz(hipStream_t st, int N, const PoseSlot* slots, const LigandDesc* ligs, const RotMeta* meta,
                         int nslots, int group, const float2* Z, int cmax, int boxmax, const float2* P2, float2* T2,
                         float invG, const float2* tw);
hipError_t launch_inv_z(hipStream_t st, int N, const RotMeta* meta, int nslots, const float2* T2,
                        unsigned long long* keys, const float2* tw);
hipError_t launch_to_kz_blocks(hipStream_t st, int N, const float2* in, float2* out, int nch);
// Atom form (grid edges 64, 96 and 128): the mid stage builds each channel's
// ligand spectrum on a kz plane as a sum of per-atom outer products on the
// matrix cores (no ligand grids, no forward stage).  At 64 it then runs as
// the column form (the inverse y and x transforms in the mid kernel, window
// rows into T2, k_inv_z last); at 96 and 128 the mid stage runs the inverse
// y transforms and writes the window's y rows of the product spectra into T2
// (at most N (N/2 + 1) N float2 per pose), and launch_inv_z runs the inverse
// x and z transforms and the argmin per (pose, y row).  P3 is the protein
// spectra in the matrix-core result layout (launch_to_mfma_blocks).
constexpr int kPipeAtoms = 2;
bool atoms_supported(int N);
// Atom form ligand-spectrum products: FP16 matrix cores on operands split
// into two FP16 parts each (hi, lo; about 2^-21 per product), FP32 matrix
// cores, or FP16 matrix cores on operands split exactly into three FP16 parts
// (hi, mid, lo) with the six partial products down to 2^-22 (an FP32 x FP32
// product to about 2^-32, below FP32's own rounding).  In the split forms the
// operands are scaled by powers of two (LigandDesc::ascale for the atom
// weights, 2^12 for the y factors) so no part is an FP16 subnormal above
// 2^-25 of the largest; the scale is taken out exactly before the inverse.
constexpr int kAtomF16 = 0, kAtomF32 = 1, kAtomF16x3 = 2;
constexpr float kAtomYScale = 4096.f;
// pad_atoms: the ligands' padded channel lists (LigandDesc::pad_atoms).
hipError_t launch_mid_atoms(hipStream_t st, int N, const PoseSlot* slots, const LigandDesc* ligs,
                            const int* pad_atoms, const float* wv, const float* wq, const AtomGrid* ag, int astride,
                            const RotMeta* meta, int nslots, int group, const float2* P3, float2* T2, float invG,
                            const float2* tw, int products);
hipError_t launch_to_mfma_blocks(hipStream_t st, int N, const float2* in, float2* out, int nch);
// Inverse plane stage variants: one workgroup per (pose, window plane) with
// block-wide LDS FFTs, or one workgroup per (pose, 4 window planes) with
// register FFTs and one LDS tile of N rows.
constexpr int kInvBlock = 0, kInvReg = 1;
bool inv_reg_supported(int N);
hipError_t launch_inv_argmin(hipStream_t st, int N, int variant, const RotMeta* meta, int nslots, int wmax,
                             const float2* T, unsigned long long* keys, const float2* tw);

hipError_t launch_rotate(hipStream_t st, int nq, int natom, const float* quats, const float* coords, float* out);

// Standalone FillLigandGrid on caller-rotated coordinates (+= semantics).
hipError_t launch_place(hipStream_t st, int nq, int natom, const float* coors, const float* mincoors, float h,
                        AtomGrid* ag);
hipError_t launch_fill_generic(hipStream_t st, int nq, int natom, const AtomGrid* ag, const int* ch_start,
                               const int* chan_atoms, const float* wv, const float* wq, int C, int X, int Y, int Z,
                               float* grid);

}  // namespace cdk
