// Plain-old-data types shared by the host engine and the gfx950 kernels.
#pragma once
#include <cstdint>

namespace cdk {

constexpr int kMaxCh = 16;       // used channels per ligand (vdW radii used + 1 elec)
constexpr int kMaxAtoms = 512;   // atoms per ligand
constexpr int kBlock = 256;      // threads per workgroup for every grid kernel
constexpr int kPrepBlock = 64;   // one wavefront per pose in the prep kernel

// One ligand as the kernels see it.  Channel c (0 <= c < nch) lists its atoms
// in chan_atoms[list_off + ch_start[c] .. list_off + ch_start[c+1]) in
// ascending atom order, which is the order FillLigandGrid.cpp:18-72 (vdW) and
// :75-119 (elec) accumulates them.  The last channel (nch-1) is the
// electrostatic one and weights atoms by charge; the others weight by
// sqrt(epsilon) (FillLigandGrid.cpp:45 and :97).
struct LigandDesc {
  int natom;
  int nch;
  int w_off;      // per-atom weights: wv[w_off + a] = sqrtf(eps_a), wq[w_off + a] = q_a
  int list_off;   // into chan_atoms
  int ch_start[kMaxCh + 1];
  int pmap[kMaxCh];  // protein channel used for ligand channel c (TranRotaConfSearchNoMini.cu:136-146)
  // The same lists with each channel padded to a multiple of 8 entries by -1
  // (no atom): channel c at pad_atoms[pad_off + pad_start[c] ..), pad_start[c]
  // a multiple of 8, pad_start[nch] the padded total.  Used by the atom form,
  // whose matrix-core steps take 8 entries at a time.
  int pad_off;
  int pad_start[kMaxCh + 1];
  // Power of two that brings the largest atom weight (|sqrt(eps)| or |q|) of
  // the ligand to [2^12, 2^13), and its inverse; the atom form's FP16 split
  // products use it so no operand part underflows.
  float ascale, aunscale;
};

// One pose = one (ligand, conformer, rotation) correlation.
struct PoseSlot {
  float q[4];      // quaternion (w, x, y, z), QuaternionUniformSampling.cpp:19-22
  int lig;         // index into the LigandDesc table
  int coord_off;   // float offset of the conformer's coordinates [natom][3]
  int pad[2];
};

// Per-pose results of the prep kernel.
struct RotMeta {
  float minc[3];   // GetMinCoors.cpp:4-29 of the rotated coordinates
  float len[3];    // max - min, TranRotaConfSearchNoMini.cu:367-372
  int box[3];      // occupied ligand box: max grid index + 2 along each axis
  int win[3];      // translation window N - int(len/h) - 2 (TranRotaConfSearchNoMini.cu:446-450)
  int valid;       // window non-empty and box inside the grid
  int pad;
};

// Trilinear placement of one rotated atom (FillLigandGrid.cpp:24-35).
struct AtomGrid {
  int ix, iy, iz, pad;
  double fx, fy, fz;
};

}  // namespace cdk
