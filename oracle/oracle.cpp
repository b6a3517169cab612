// ORACLE - test infrastructure only.
//
// A CPU restatement of the reference's translational-rotational search, used
// by tests/, by the smoke check and by the CPU-baseline leg of bench.py as the
// checker.  Nothing in the product links or calls this file.
//
// Functions and the reference code each follows:
//   orc_rotate                    Rotate.cpp:4-27, QuaternionMultiply.cpp:3-11
//   orc_min_coors / orc_max_coors GetMinCoors.cpp:4-29, GetMaxCoors.cpp:3-27
//   orc_idx_of_atoms_for_vdw_radius GetIdxOfAtomsForVdwRadius.cpp:4-35
//   orc_fill_ligand_grid          FillLigandGrid.cpp:7-120
//   orc_quaternion_uniform_sampling QuaternionUniformSampling.cpp:5-24
//   orc_filter_quaternions        FilterQuaternions.cpp:11-83 (no-shuffle branch)
//   orc_correlate_window_{f32,f64} cpuNaive.cpp:184-216 (dense direct
//                                 correlation, FP32 or FP64 accumulation, order
//                                 c, m, n, o), with the memcpy size defect of
//                                 cpuNaive.cpp:69-77 fixed
//   orc_argmin_first              TranRotaConfSearchNoMini.cu:441-464 (strict
//                                 '<' scan, x-major, first minimum wins)
//   orc_search                    NoMini:352-372, 390-409, 441-464 per pose
//   orc_sparse_energy_f64         the same energy as a per-atom trilinear
//                                 gather, an independent FP64 formulation
//   orc_sparse_search_f64         that gather over the whole window: FP64
//                                 argmin of many poses on worker threads
//   orc_build_grids               the protein potential grids from atoms (the
//                                 synthetic recipe's definition), brute force
//
// Pinning: see tests/test_oracle.py (golden vectors from the compiled
// reference sources, closed forms, invariants, brute force, planted faults).
// Build: oracle/Makefile (g++, -ffp-contract=off so float/double expressions
// round exactly as written).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <thread>
#include <vector>

// ---- A3: Rotate = Im(q (x) (0,p) (x) conj(q)), Hamilton product term order.
static void qmul(const float* l, const float* r, float* o) {
  o[0] = l[0] * r[0] - l[1] * r[1] - l[2] * r[2] - l[3] * r[3];
  o[1] = l[0] * r[1] + l[1] * r[0] + l[2] * r[3] - l[3] * r[2];
  o[2] = l[0] * r[2] - l[1] * r[3] + l[2] * r[0] + l[3] * r[1];
  o[3] = l[0] * r[3] + l[1] * r[2] - l[2] * r[1] + l[3] * r[0];
}

extern "C" void orc_rotate(const float* q, const float* p, float* out) {
  const float qc[4] = {q[0], -q[1], -q[2], -q[3]};
  const float qp[4] = {0.0f, p[0], p[1], p[2]};
  float t1[4], t2[4];
  qmul(q, qp, t1);
  qmul(t1, qc, t2);
  out[0] = t2[1];
  out[1] = t2[2];
  out[2] = t2[3];
}

extern "C" void orc_rotate_batch(int nq, int natom, const float* quats, const float* coords, float* out) {
  for (int i = 0; i < nq; ++i)
    for (int a = 0; a < natom; ++a) orc_rotate(quats + i * 4, coords + a * 3, out + ((size_t)i * natom + a) * 3);
}

// ---- A4: first-occurrence strict min / max per axis.
extern "C" void orc_min_coors(int nq, int natom, const float* c, float* out) {
  for (int i = 0; i < nq; ++i) {
    const float* p = c + (size_t)i * natom * 3;
    for (int k = 0; k < 3; ++k) out[i * 3 + k] = p[k];
    for (int a = 1; a < natom; ++a)
      for (int k = 0; k < 3; ++k)
        if (p[a * 3 + k] < out[i * 3 + k]) out[i * 3 + k] = p[a * 3 + k];
  }
}

extern "C" void orc_max_coors(int nq, int natom, const float* c, float* out) {
  for (int i = 0; i < nq; ++i) {
    const float* p = c + (size_t)i * natom * 3;
    for (int k = 0; k < 3; ++k) out[i * 3 + k] = p[k];
    for (int a = 1; a < natom; ++a)
      for (int k = 0; k < 3; ++k)
        if (p[a * 3 + k] > out[i * 3 + k]) out[i * 3 + k] = p[a * 3 + k];
  }
}

// ---- A5: nearest probe radius per atom (first wins ties), used radii.
extern "C" void orc_idx_of_atoms_for_vdw_radius(int natom, const float* ar, int nv, const float* gr, int* num_used,
                                     int* idx_used, int* atom_radius_index) {
  std::vector<int> cnt(nv, 0);
  for (int i = 0; i < natom; ++i) {
    int idx = 0;
    double diff = std::fabs(ar[i] - gr[0]);
    for (int j = 1; j < nv; ++j)
      if (std::fabs(ar[i] - gr[j]) < diff) {
        idx = j;
        diff = std::fabs(ar[i] - gr[j]);
      }
    atom_radius_index[i] = idx;
    cnt[idx]++;
  }
  int u = 0;
  for (int j = 0; j < nv; ++j)
    if (cnt[j]) idx_used[u++] = j;
  *num_used = u;
}

// ---- A7: trilinear spreading, += into a caller-zeroed [q][C][X][Y][Z].
static void spread_one(float* g, int Y, int Z, double x, double y, double z, float mx, float my, float mz,
                       float spacing, float w) {
  const int xi = int((x - mx) / spacing), yi = int((y - my) / spacing), zi = int((z - mz) / spacing);
  const double xr = ((x - mx) - xi * spacing) / spacing;
  const double yr = ((y - my) - yi * spacing) / spacing;
  const double zr = ((z - mz) - zi * spacing) / spacing;
  const int dx[8] = {0, 0, 0, 0, 1, 1, 1, 1}, dy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, dz[8] = {0, 1, 0, 1, 0, 1, 0, 1};
  for (int k = 0; k < 8; ++k) {
    const double fx = dx[k] ? xr : (1 - xr), fy = dy[k] ? yr : (1 - yr), fz = dz[k] ? zr : (1 - zr);
    const int idx = ((xi + dx[k]) * Y + yi + dy[k]) * Z + zi + dz[k];
    g[idx] += w * fx * fy * fz;
  }
}

extern "C" void orc_fill_ligand_grid(int nq, int natom, const float* coors, const float* mincoors, const float* charge,
                          const float* eps, int num_vdw_used, const int* idx_vdw_used, const int* radius_offsets,
                          const int* radius_atoms, int X, int Y, int Z, float spacing, float* grid) {
  const size_t G = (size_t)X * Y * Z;
  const int C = num_vdw_used + 1;
  for (int i = 0; i < nq; ++i)
    for (int j = 0; j < num_vdw_used; ++j) {
      const int r = idx_vdw_used[j];
      for (int l = radius_offsets[r]; l < radius_offsets[r + 1]; ++l) {
        const int m = radius_atoms[l];
        const float* p = coors + ((size_t)i * natom + m) * 3;
        spread_one(grid + ((size_t)i * C + j) * G, Y, Z, p[0], p[1], p[2], mincoors[i * 3], mincoors[i * 3 + 1],
                   mincoors[i * 3 + 2], spacing, std::sqrt(eps[m]));
      }
    }
  for (int i = 0; i < nq; ++i)
    for (int m = 0; m < natom; ++m) {
      const float* p = coors + ((size_t)i * natom + m) * 3;
      spread_one(grid + ((size_t)i * C + num_vdw_used) * G, Y, Z, p[0], p[1], p[2], mincoors[i * 3],
                 mincoors[i * 3 + 1], mincoors[i * 3 + 2], spacing, charge[m]);
    }
}

// ---- A1: Shoemake sampling with the standard mt19937_64 / uniform_real_distribution.
// ncall successive calls of n each on one generator (per-conformer redraw).
extern "C" void orc_quaternion_uniform_sampling(uint64_t seed, int ncall, int n, float* out) {
  std::mt19937_64 gen(seed);
  for (int c = 0; c < ncall; ++c) {
    std::uniform_real_distribution<double> u(0, 1);
    float* o = out + (size_t)c * n * 4;
    for (int i = 0; i < n; ++i) {
      const double s = u(gen);
      const double s1 = std::sqrt(1 - s), s2 = std::sqrt(s);
      const double t1 = 2 * M_PI * u(gen), t2 = 2 * M_PI * u(gen);
      o[i * 4 + 0] = (float)(std::cos(t2) * s2);
      o[i * 4 + 1] = (float)(std::sin(t1) * s1);
      o[i * 4 + 2] = (float)(std::cos(t1) * s1);
      o[i * 4 + 3] = (float)(std::sin(t2) * s2);
    }
  }
}

// ---- A2: filter by rotated extent (strict '<'), no-shuffle branch.
extern "C" int orc_filter_quaternions(const float* coor, int natom, int nq, const float* quats, float lx, float ly, float lz,
                           int* kept) {
  std::vector<float> rot((size_t)natom * 3), mn(3), mx(3);
  int n = 0;
  for (int i = 0; i < nq; ++i) {
    orc_rotate_batch(1, natom, quats + i * 4, coor, rot.data());
    orc_min_coors(1, natom, rot.data(), mn.data());
    orc_max_coors(1, natom, rot.data(), mx.data());
    const float l0 = mx[0] - mn[0], l1 = mx[1] - mn[1], l2 = mx[2] - mn[2];
    if (l0 < lx && l1 < ly && l2 < lz) kept[n++] = i;
  }
  return n;
}

// ---- A13: first strict minimum in x-major order.
extern "C" void orc_argmin_first(const float* E, int wx, int wy, int wz, float* emin, int* idx3) {
  float m = INFINITY;
  idx3[0] = idx3[1] = idx3[2] = -1;
  for (int i = 0; i < wx; ++i)
    for (int j = 0; j < wy; ++j)
      for (int k = 0; k < wz; ++k) {
        const float e = E[((size_t)i * wy + j) * wz + k];
        if (e < m) {
          m = e;
          idx3[0] = i;
          idx3[1] = j;
          idx3[2] = k;
        }
      }
  *emin = m;
}

static int g_threads = 0;
// Worker threads of the correlation loops (0 = hardware concurrency).
extern "C" void orc_set_threads(int n) { g_threads = n; }
static int n_threads(int cap) {
  const int hw = g_threads > 0 ? g_threads : (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(cap, hw));
}

// ---- A15: dense direct correlation over the translation window.
//   E[i][j][k] = sum_c sum_m sum_n sum_o L[c][m][n][o] * P[c][m+i][n+j][o+k]
// accumulated in T in the order c, m, n, o.  Cells with L == 0 add an exact
// +-0 product, which never changes a sum that starts at +0, so they are
// skipped; the result is bit-identical to the full loop.
template <class T>
static void correlate_window(const float* P, int N, const float* L, int C, int bx, int by, int bz, int wx, int wy,
                             int wz, T* E) {
  struct Cell {
    int c, m, n, o;
    T v;
  };
  std::vector<Cell> cells;
  for (int c = 0; c < C; ++c)
    for (int m = 0; m < bx; ++m)
      for (int n = 0; n < by; ++n)
        for (int o = 0; o < bz; ++o) {
          const float v = L[(((size_t)c * bx + m) * by + n) * bz + o];
          if (v != 0.0f) cells.push_back({c, m, n, o, (T)v});
        }
  const size_t G = (size_t)N * N * N;
  auto work = [&](int i0, int i1) {
    for (int i = i0; i < i1; ++i)
      for (int j = 0; j < wy; ++j)
        for (int k = 0; k < wz; ++k) {
          T e = 0;
          for (const Cell& cl : cells) {
            const T p = (T)P[(size_t)cl.c * G + ((size_t)(cl.m + i) * N + (cl.n + j)) * N + (cl.o + k)];
            const T prod = cl.v * p;
            e = e + prod;
          }
          E[((size_t)i * wy + j) * wz + k] = e;
        }
  };
  const int nt = n_threads(wx);
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) th.emplace_back(work, wx * t / nt, wx * (t + 1) / nt);
  for (auto& t : th) t.join();
}

// The literal cpuNaive.cpp:194-216 loop nest: every cell of the ligand box,
// zeros included, FP32 accumulation in c, m, n, o order.  Used as the timed
// CPU baseline (bench.py); results equal correlate_window<float>.
extern "C" void orc_correlate_window_full_f32(const float* P, int N, const float* L, int C, int bx, int by, int bz,
                                              int wx, int wy, int wz, float* E) {
  const size_t G = (size_t)N * N * N;
  auto work = [&](int i0, int i1) {
    for (int i = i0; i < i1; ++i)
      for (int j = 0; j < wy; ++j)
        for (int k = 0; k < wz; ++k) {
          float e = 0.0f;
          for (int c = 0; c < C; ++c)
            for (int m = 0; m < bx; ++m)
              for (int n = 0; n < by; ++n) {
                const float* lp = L + (((size_t)c * bx + m) * by + n) * bz;
                const float* pp = P + (size_t)c * G + ((size_t)(m + i) * N + (n + j)) * N + k;
                for (int o = 0; o < bz; ++o) e = e + lp[o] * pp[o];
              }
          E[((size_t)i * wy + j) * wz + k] = e;
        }
  };
  const int nt = n_threads(wx);
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) th.emplace_back(work, wx * t / nt, wx * (t + 1) / nt);
  for (auto& t : th) t.join();
}

extern "C" void orc_correlate_window_f32(const float* P, int N, const float* L, int C, int bx, int by, int bz, int wx, int wy,
                              int wz, float* E) {
  correlate_window<float>(P, N, L, C, bx, by, bz, wx, wy, wz, E);
}

extern "C" void orc_correlate_window_f64(const float* P, int N, const float* L, int C, int bx, int by, int bz, int wx, int wy,
                              int wz, double* E) {
  correlate_window<double>(P, N, L, C, bx, by, bz, wx, wy, wz, E);
}

// ---- Per-pose preparation (NoMini:352-372, 390-409): rotate, AABB, window,
// ligand box grid [C][bx][by][bz] with FillLigandGrid arithmetic.  The box is
// the larger of cpuNaive's int(len/h)+2 (cpuNaive.cpp:103-105) and the
// highest occupied index + 2, so every non-zero cell is inside it.
// Channels: vdW channels in idx_used order, elec last (NoMini:136-146).
struct Prep {
  std::vector<float> rot;
  float mn[3], mx[3], len[3];
  int win[3], box[3];
  std::vector<float> L;
};

static void prepare(int natom, const float* coords, const float* q, const float* charge, const float* eps,
                    int num_used, const int* idx_used, const int* atom_radius_index, int N, float h, Prep& p) {
  p.rot.resize((size_t)natom * 3);
  orc_rotate_batch(1, natom, q, coords, p.rot.data());
  orc_min_coors(1, natom, p.rot.data(), p.mn);
  orc_max_coors(1, natom, p.rot.data(), p.mx);
  int imax[3] = {0, 0, 0};
  for (int a = 0; a < natom; ++a)
    for (int k = 0; k < 3; ++k) imax[k] = std::max(imax[k], int((p.rot[a * 3 + k] - (double)p.mn[k]) / h));
  for (int k = 0; k < 3; ++k) {
    p.len[k] = p.mx[k] - p.mn[k];
    const int il = int(p.len[k] / h);
    p.win[k] = N - il - 2;
    p.box[k] = std::max(il + 2, imax[k] + 2);
  }
  const int C = num_used + 1;
  p.L.assign((size_t)C * p.box[0] * p.box[1] * p.box[2], 0.0f);
  std::vector<int> offs(64 + 1, 0), atoms;
  // CSR of atoms per probe radius (ascending atom order).
  int nv = 0;
  for (int a = 0; a < natom; ++a) nv = std::max(nv, atom_radius_index[a] + 1);
  offs.assign(nv + 1, 0);
  for (int r = 0; r < nv; ++r) {
    offs[r] = (int)atoms.size();
    for (int a = 0; a < natom; ++a)
      if (atom_radius_index[a] == r) atoms.push_back(a);
  }
  offs[nv] = (int)atoms.size();
  orc_fill_ligand_grid(1, natom, p.rot.data(), p.mn, charge, eps, num_used, idx_used, offs.data(), atoms.data(),
                       p.box[0], p.box[1], p.box[2], h, p.L.data());
}

// One pose: oracle window energies (FP32 or FP64) and the reference argmin.
// Pused: the protein channels the ligand uses, [C][N^3].  E_out (optional)
// receives the window energies [wx][wy][wz] as double.
extern "C" int orc_search_pose(const float* Pused, int N, float h, int natom, const float* coords, const float* q,
                    const float* charge, const float* eps, int num_used, const int* idx_used,
                    const int* atom_radius_index, int use_double, float* emin, int* idx3, float* minc3, int* win3,
                    double* E_out) {
  Prep p;
  prepare(natom, coords, q, charge, eps, num_used, idx_used, atom_radius_index, N, h, p);
  const int C = num_used + 1;
  for (int k = 0; k < 3; ++k) {
    minc3[k] = p.mn[k];
    win3[k] = p.win[k];
  }
  if (p.win[0] <= 0 || p.win[1] <= 0 || p.win[2] <= 0 || p.box[0] > N || p.box[1] > N || p.box[2] > N) {
    *emin = INFINITY;
    idx3[0] = idx3[1] = idx3[2] = -1;
    return 0;
  }
  const size_t W = (size_t)p.win[0] * p.win[1] * p.win[2];
  std::vector<float> Ef(W);
  if (use_double) {
    std::vector<double> Ed(W);
    correlate_window<double>(Pused, N, p.L.data(), C, p.box[0], p.box[1], p.box[2], p.win[0], p.win[1], p.win[2],
                             Ed.data());
    for (size_t i = 0; i < W; ++i) Ef[i] = (float)Ed[i];
    if (E_out) std::copy(Ed.begin(), Ed.end(), E_out);
    // FP64 argmin on the doubles themselves.
    double m = INFINITY;
    idx3[0] = idx3[1] = idx3[2] = -1;
    for (int i = 0; i < p.win[0]; ++i)
      for (int j = 0; j < p.win[1]; ++j)
        for (int k = 0; k < p.win[2]; ++k) {
          const double e = Ed[((size_t)i * p.win[1] + j) * p.win[2] + k];
          if (e < m) {
            m = e;
            idx3[0] = i;
            idx3[1] = j;
            idx3[2] = k;
          }
        }
    *emin = (float)m;
    return 1;
  }
  correlate_window<float>(Pused, N, p.L.data(), C, p.box[0], p.box[1], p.box[2], p.win[0], p.win[1], p.win[2],
                          Ef.data());
  if (E_out)
    for (size_t i = 0; i < W; ++i) E_out[i] = Ef[i];
  orc_argmin_first(Ef.data(), p.win[0], p.win[1], p.win[2], emin, idx3);
  return 1;
}

// Ligand box grid of one pose (for tests that inspect the spread).
extern "C" int orc_pose_box(int N, float h, int natom, const float* coords, const float* q, const float* charge,
                 const float* eps, int num_used, const int* idx_used, const int* atom_radius_index, int* box3,
                 int* win3, float* minc3, float* L, size_t Lcap) {
  Prep p;
  prepare(natom, coords, q, charge, eps, num_used, idx_used, atom_radius_index, N, h, p);
  for (int k = 0; k < 3; ++k) {
    box3[k] = p.box[k];
    win3[k] = p.win[k];
    minc3[k] = p.mn[k];
  }
  if (L && p.L.size() <= Lcap) std::copy(p.L.begin(), p.L.end(), L);
  return (int)p.L.size();
}

// ---- Independent FP64 formulation: E(t) = sum_atoms w * trilinear gather of
// the atom's channel at (i, j, k) + t, using the same grid placement.
extern "C" double orc_sparse_energy_f64(const float* Pused, int N, float h, int natom, const float* rot, const float* minc3,
                             const float* charge, const float* eps, int num_used, const int* atom_channel, int ti,
                             int tj, int tk) {
  const size_t G = (size_t)N * N * N;
  double e = 0;
  for (int a = 0; a < natom; ++a) {
    int idx[3];
    double f[3];
    for (int k = 0; k < 3; ++k) {
      const double d = (double)rot[a * 3 + k] - (double)minc3[k];
      idx[k] = (int)(d / h);
      f[k] = (d - (double)((float)idx[k] * h)) / h;
    }
    for (int pass = 0; pass < 2; ++pass) {
      const int c = pass == 0 ? atom_channel[a] : num_used;
      if (c < 0) continue;
      const double w = pass == 0 ? (double)std::sqrt(eps[a]) : (double)charge[a];
      double s = 0;
      for (int cx = 0; cx < 2; ++cx)
        for (int cy = 0; cy < 2; ++cy)
          for (int cz = 0; cz < 2; ++cz) {
            const double W = (cx ? f[0] : 1 - f[0]) * (cy ? f[1] : 1 - f[1]) * (cz ? f[2] : 1 - f[2]);
            const size_t gi = ((size_t)(idx[0] + cx + ti) * N + (idx[1] + cy + tj)) * N + (idx[2] + cz + tk);
            s += W * (double)Pused[(size_t)c * G + gi];
          }
      e += w * s;
    }
  }
  return e;
}


// ---- Full-window FP64 search by the same per-atom trilinear gather
// (NoMini:441-464 window and first-minimum rule over E64(t) of
// orc_sparse_energy_f64), for many poses on worker threads.  Per pose p:
// rotate conformer conf[p] by quats[p] (Rotate.cpp:4-27), AABB
// (GetMinCoors.cpp:4-29, GetMaxCoors.cpp:3-27), window N - int(len/h) - 2;
// then over every window translation E64(t) = sum_a w_a sum_corners W P[..+t]
// with the corner weights and grid offsets formed once per atom.  Outputs:
// valid[p], emin[p], idx[p][3] (first minimum in x-major order); and, when
// t_at is given, E64 and the absolute scale S = sum |w W P| at t_at[p] (the
// caller's translation, e.g. the GPU's argmin).  full = 0 skips the window
// search (only the values at t_at).
extern "C" void orc_sparse_search_f64(const float* Pused, int N, float h, int natom, const float* coords,
                                      const int* conf, const float* quats, int npose, const float* charge,
                                      const float* eps, int num_used, const int* atom_channel, const int* t_at,
                                      int full, int* valid, double* emin, int* idx, double* e_at, double* s_at) {
  const size_t G = (size_t)N * N * N;
  auto one = [&](int p) {
    std::vector<float> rot((size_t)natom * 3);
    orc_rotate_batch(1, natom, quats + (size_t)p * 4, coords + (size_t)conf[p] * natom * 3, rot.data());
    float mn[3], mx[3];
    orc_min_coors(1, natom, rot.data(), mn);
    orc_max_coors(1, natom, rot.data(), mx);
    int win[3], imax[3] = {0, 0, 0};
    for (int a = 0; a < natom; ++a)
      for (int k = 0; k < 3; ++k) imax[k] = std::max(imax[k], int((rot[a * 3 + k] - (double)mn[k]) / h));
    bool ok = true;
    for (int k = 0; k < 3; ++k) {
      const float len = mx[k] - mn[k];
      win[k] = N - int(len / h) - 2;
      ok = ok && win[k] > 0 && imax[k] + 2 <= N;
    }
    valid[p] = ok;
    emin[p] = INFINITY;
    idx[p * 3] = idx[p * 3 + 1] = idx[p * 3 + 2] = -1;
    if (e_at) e_at[p] = NAN;
    if (s_at) s_at[p] = NAN;
    if (!ok) return;
    // per (atom, channel pass, corner): weight w W and grid offset
    std::vector<double> wt;
    std::vector<size_t> off;
    wt.reserve((size_t)natom * 16);
    off.reserve((size_t)natom * 16);
    for (int a = 0; a < natom; ++a) {
      int id[3];
      double f[3];
      for (int k = 0; k < 3; ++k) {
        const double d = (double)rot[a * 3 + k] - (double)mn[k];
        id[k] = (int)(d / h);
        f[k] = (d - (double)((float)id[k] * h)) / h;
      }
      for (int pass = 0; pass < 2; ++pass) {
        const int c = pass == 0 ? atom_channel[a] : num_used;
        if (c < 0) continue;
        const double w = pass == 0 ? (double)std::sqrt(eps[a]) : (double)charge[a];
        for (int cx = 0; cx < 2; ++cx)
          for (int cy = 0; cy < 2; ++cy)
            for (int cz = 0; cz < 2; ++cz) {
              const double W = (cx ? f[0] : 1 - f[0]) * (cy ? f[1] : 1 - f[1]) * (cz ? f[2] : 1 - f[2]);
              wt.push_back(w * W);
              off.push_back((size_t)c * G + ((size_t)(id[0] + cx) * N + (id[1] + cy)) * N + (id[2] + cz));
            }
      }
    }
    const size_t nt = wt.size();
    double m = INFINITY;
    // one window row (i, j) at a time: every term adds to all of the row's
    // translations, in term order (each E(t) sums its terms in the same order
    // as a per-translation loop would)
    std::vector<double> row(win[2]);
    for (int i = 0; i < (full ? win[0] : 0); ++i)
      for (int j = 0; j < win[1]; ++j) {
        std::fill(row.begin(), row.end(), 0.0);
        const size_t to = ((size_t)i * N + j) * N;
        for (size_t u = 0; u < nt; ++u) {
          const float* pp = Pused + off[u] + to;
          const double w = wt[u];
          for (int k = 0; k < win[2]; ++k) row[k] += w * (double)pp[k];
        }
        for (int k = 0; k < win[2]; ++k)
          if (row[k] < m) {
            m = row[k];
            idx[p * 3] = i;
            idx[p * 3 + 1] = j;
            idx[p * 3 + 2] = k;
          }
      }
    emin[p] = m;
    if (t_at) {
      const int* t = t_at + (size_t)p * 3;
      if (t[0] >= 0 && t[1] >= 0 && t[2] >= 0 && t[0] < win[0] && t[1] < win[1] && t[2] < win[2]) {
        const size_t to = ((size_t)t[0] * N + t[1]) * N + t[2];
        double e = 0, sabs = 0;
        for (size_t u = 0; u < nt; ++u) {
          const double v = wt[u] * (double)Pused[off[u] + to];
          e += v;
          sabs += std::fabs(v);
        }
        if (e_at) e_at[p] = e;
        if (s_at) s_at[p] = sabs;
      }
    }
  };
  const int nth = n_threads(npose);
  std::vector<std::thread> th;
  for (int t = 0; t < nth; ++t)
    th.emplace_back([&, t] {
      for (int p = t; p < npose; p += nth) one(p);
    });
  for (auto& t : th) t.join();
}

// ---- Protein potential grids from atoms (the grid builder's definition):
// LJ channel p: sum_i sqrt(eps_i) ((R/d)^12 - 2 (R/d)^6), R = r_i + probe_p,
// d <= rc_lj, capped at +lj_cap; Coulomb: sum_i 332.0636 q_i / (4 d^2),
// d <= rc_el, clamped to +-el_cap; d^2 floored at 1e-4.  Brute force over all
// atoms in ascending index order, FP64.  atoms [natom][6] = x, y, z, q, r, eps.
extern "C" void orc_build_grids(int natom, const double* atoms, int nprobe, const double* probe, int n,
                                const double* gmin, double h, double rc_lj, double rc_el, double lj_cap,
                                double el_cap, float* out) {
  const size_t G = (size_t)n * n * n;
  std::vector<double> lj(nprobe);
  for (int ix = 0; ix < n; ++ix)
    for (int iy = 0; iy < n; ++iy)
      for (int iz = 0; iz < n; ++iz) {
        const double gx = gmin[0] + ix * h, gy = gmin[1] + iy * h, gz = gmin[2] + iz * h;
        std::fill(lj.begin(), lj.end(), 0.0);
        double el = 0.0;
        for (int i = 0; i < natom; ++i) {
          const double* a = atoms + (size_t)i * 6;
          const double dx = a[0] - gx, dy = a[1] - gy, dz = a[2] - gz;
          double d2 = dx * dx + dy * dy + dz * dz;
          if (d2 > rc_el * rc_el) continue;
          d2 = std::max(d2, 1e-4);
          el += 332.0636 * a[3] / (4.0 * d2);
          if (d2 > rc_lj * rc_lj) continue;
          const double inv2 = 1.0 / d2;
          const double se = std::sqrt(a[5]);
          for (int p = 0; p < nprobe; ++p) {
            const double R = a[4] + probe[p];
            const double s2 = R * R * inv2;
            const double s6 = s2 * s2 * s2;
            lj[p] += se * (s6 * s6 - 2.0 * s6);
          }
        }
        const size_t g = ((size_t)ix * n + iy) * n + iz;
        for (int p = 0; p < nprobe; ++p) out[(size_t)p * G + g] = (float)std::min(lj[p], lj_cap);
        out[(size_t)nprobe * G + g] = (float)std::max(-el_cap, std::min(el, el_cap));
      }
}

// ---- Random-torsion conformers (GeneRandomConformations.cpp:36-47 with a
// counter-based stream and a steric filter in place of OpenMM): every rotor
// (a, b, c, d) in order is set to theta = -pi + 2 pi u(seed, trial, rotor);
// the atoms flagged in moving[r] turn about b->c through c by theta - phi,
// phi the current (IUPAC) dihedral.  ok[t] = 1 when no non-excluded pair is closer
// than clash_dist.  FP64, plain loops.
static double orc_torsion_uniform(uint64_t seed, int trial, int rotor) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)trial * 0x100000001ull + (uint64_t)rotor + 1ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

extern "C" void orc_conformers(int natom, const double* coords, int nrotor, const int* rotors,
                               const unsigned char* moving, const unsigned char* exclude, int ntrial, uint64_t seed,
                               double clash_dist, double* out, int* ok) {
  const double kPi = 3.14159265358979323846;
  for (int t = 0; t < ntrial; ++t) {
    double* x = out + (size_t)t * natom * 3;
    std::memcpy(x, coords, sizeof(double) * natom * 3);
    for (int r = 0; r < nrotor; ++r) {
      const int a = rotors[r * 4 + 0], b = rotors[r * 4 + 1], c = rotors[r * 4 + 2], d = rotors[r * 4 + 3];
      double b1[3], b2[3], b3[3];
      for (int k = 0; k < 3; ++k) {
        b1[k] = x[b * 3 + k] - x[a * 3 + k];
        b2[k] = x[c * 3 + k] - x[b * 3 + k];
        b3[k] = x[d * 3 + k] - x[c * 3 + k];
      }
      const double n1x = b1[1] * b2[2] - b1[2] * b2[1], n1y = b1[2] * b2[0] - b1[0] * b2[2],
                   n1z = b1[0] * b2[1] - b1[1] * b2[0];
      const double n2x = b2[1] * b3[2] - b2[2] * b3[1], n2y = b2[2] * b3[0] - b2[0] * b3[2],
                   n2z = b2[0] * b3[1] - b2[1] * b3[0];
      const double l2 = std::sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
      // phi = atan2(b2^ . (n1 x n2), n1 . n2), the IUPAC dihedral
      const double cx_ = n1y * n2z - n1z * n2y, cy_ = n1z * n2x - n1x * n2z, cz_ = n1x * n2y - n1y * n2x;
      const double yv = (cx_ * b2[0] + cy_ * b2[1] + cz_ * b2[2]) / l2;
      const double xv = n1x * n2x + n1y * n2y + n1z * n2z;
      const double phi = std::atan2(yv, xv);
      const double theta = -kPi + 2.0 * kPi * orc_torsion_uniform(seed, t, r);
      const double ang = theta - phi;
      const double ux = b2[0] / l2, uy = b2[1] / l2, uz = b2[2] / l2;
      const double ca = std::cos(ang), sa = std::sin(ang);
      const double cx = x[c * 3 + 0], cy = x[c * 3 + 1], cz = x[c * 3 + 2];
      for (int i = 0; i < natom; ++i) {
        if (!moving[(size_t)r * natom + i]) continue;
        const double vx = x[i * 3 + 0] - cx, vy = x[i * 3 + 1] - cy, vz = x[i * 3 + 2] - cz;
        const double kd = ux * vx + uy * vy + uz * vz;
        const double kx = uy * vz - uz * vy, ky = uz * vx - ux * vz, kz = ux * vy - uy * vx;
        x[i * 3 + 0] = cx + vx * ca + kx * sa + ux * kd * (1.0 - ca);
        x[i * 3 + 1] = cy + vy * ca + ky * sa + uy * kd * (1.0 - ca);
        x[i * 3 + 2] = cz + vz * ca + kz * sa + uz * kd * (1.0 - ca);
      }
    }
    int bad = 0;
    for (int i = 0; i < natom && !bad; ++i)
      for (int j = i + 1; j < natom; ++j) {
        if (exclude[(size_t)i * natom + j]) continue;
        const double dx = x[i * 3 + 0] - x[j * 3 + 0], dy = x[i * 3 + 1] - x[j * 3 + 1], dz = x[i * 3 + 2] - x[j * 3 + 2];
        if (dx * dx + dy * dy + dz * dz < clash_dist * clash_dist) {
          bad = 1;
          break;
        }
      }
    ok[t] = bad ? 0 : 1;
  }
}
