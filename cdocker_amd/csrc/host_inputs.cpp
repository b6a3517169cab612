// Host-side input producers: seeded synthetic workloads, the reference's
// quaternion sampling and channel assignment, and the grid text format.
#include "host_inputs.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <thread>

#include "engine.hpp"

namespace cdk {

// ---------------------------------------------------------------------------
// Portable seeded stream: splitmix64 seeding of xoshiro256**, 53-bit uniform
// doubles, Box-Muller normals.  Same numbers on every platform.
// ---------------------------------------------------------------------------
static inline uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

Xoshiro::Xoshiro(uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
}

uint64_t Xoshiro::next() {
  auto rotl = [](uint64_t v, int k) { return (v << k) | (v >> (64 - k)); };
  const uint64_t r = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return r;
}

double Xoshiro::uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }

double Xoshiro::normal() {
  const double two_pi = 6.283185307179586476925286766559;
  const double u1 = 1.0 - uniform();  // (0, 1]
  const double u2 = uniform();
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2);
}

const float kProbeRadii[8] = {1.2f, 1.5f, 1.7f, 1.8f, 1.9f, 2.0f, 2.1f, 2.2f};

// ---------------------------------------------------------------------------
// QuaternionUniformSampling.cpp:5-24 (Shoemake), same generator and
// distribution objects, so the same seed gives the same floats.
// ---------------------------------------------------------------------------
void quaternion_uniform_sampling(std::mt19937_64& gen, int n, float* out) {
  std::uniform_real_distribution<double> udist(0, 1);
  const double two_pi = 2 * M_PI;
  for (int i = 0; i < n; ++i) {
    const double s = udist(gen);
    const double sigma1 = std::sqrt(1 - s);
    const double sigma2 = std::sqrt(s);
    const double theta1 = two_pi * udist(gen);
    const double theta2 = two_pi * udist(gen);
    out[i * 4 + 0] = (float)(std::cos(theta2) * sigma2);
    out[i * 4 + 1] = (float)(std::sin(theta1) * sigma1);
    out[i * 4 + 2] = (float)(std::cos(theta1) * sigma1);
    out[i * 4 + 3] = (float)(std::sin(theta2) * sigma2);
  }
}

void rotate_host(int nq, int natom, const float* quats, const float* coords, float* out) {
  // q (x) (0, p) (x) conj(q), Hamilton products with QuaternionMultiply.cpp:7-10's
  // term order, the zero terms kept (they decide signed zeros)
  for (int i = 0; i < nq; ++i) {
    const float* q = quats + (size_t)i * 4;
    const float c0 = q[0], c1 = -q[1], c2 = -q[2], c3 = -q[3];
    for (int a = 0; a < natom; ++a) {
      const float px = coords[a * 3], py = coords[a * 3 + 1], pz = coords[a * 3 + 2];
      const float zero = 0.0f;
      const float t0 = q[0] * zero - q[1] * px - q[2] * py - q[3] * pz;
      const float t1 = q[0] * px + q[1] * zero + q[2] * pz - q[3] * py;
      const float t2 = q[0] * py - q[1] * pz + q[2] * zero + q[3] * px;
      const float t3 = q[0] * pz + q[1] * py - q[2] * px + q[3] * zero;
      float* o = out + ((size_t)i * natom + a) * 3;
      o[0] = t0 * c1 + t1 * c0 + t2 * c3 - t3 * c2;
      o[1] = t0 * c2 - t1 * c3 + t2 * c0 + t3 * c1;
      o[2] = t0 * c3 + t1 * c2 - t2 * c1 + t3 * c0;
    }
  }
}

void min_max_coors(int nq, int natom, const float* coors, float* mincoors, float* maxcoors) {
  for (int i = 0; i < nq; ++i) {
    const float* c = coors + (size_t)i * natom * 3;
    for (int k = 0; k < 3; ++k) {
      float mn = c[k], mx = c[k];
      for (int a = 1; a < natom; ++a) {
        const float v = c[a * 3 + k];
        if (v < mn) mn = v;
        if (v > mx) mx = v;
      }
      if (mincoors) mincoors[i * 3 + k] = mn;
      if (maxcoors) maxcoors[i * 3 + k] = mx;
    }
  }
}

// ---------------------------------------------------------------------------
// Pseudo-protein potential grids.
//   LJ channel p at grid point g:  sum_i sqrt(eps_i) [(R/d)^12 - 2 (R/d)^6],
//     R = r_i + r_p (sum of radii), d >= 0.01 A, cutoff 8 A, total capped at +3.
//   Coulomb channel:  sum_i 332.0636 q_i / (4 d * d), d >= 0.01 A, cutoff 12 A,
//     capped to [-20, +20].
// The ligand energy is grid * sqrt(eps_ligand) (vdW) and grid * q (elec),
// the units the reference's grid forces use (AddGridForcesToOpenMMSystem.cpp:54,100).
// ---------------------------------------------------------------------------
void synth_protein_atoms(uint64_t seed, int natom, float box, double* xyz, double* charge, double* epsilon,
                         double* radius) {
  Xoshiro rng(seed);
  for (int i = 0; i < natom; ++i) {
    xyz[i * 3 + 0] = rng.uniform() * box;
    xyz[i * 3 + 1] = rng.uniform() * box;
    xyz[i * 3 + 2] = rng.uniform() * box;
    charge[i] = 0.3 * rng.normal();
    radius[i] = kProbeRadii[(int)(rng.uniform() * 8) & 7];
    epsilon[i] = 0.05 + 0.20 * rng.uniform();
  }
}

void synth_protein(uint64_t seed, int natom, float box, int n, float h, float* values, float* radii, float* mid3,
                   float* len3) {
  std::vector<double> xyz((size_t)natom * 3), q(natom), r(natom), eps(natom);
  synth_protein_atoms(seed, natom, box, xyz.data(), q.data(), eps.data(), r.data());
  std::vector<double> px(natom), py(natom), pz(natom), se(natom);
  for (int i = 0; i < natom; ++i) {
    px[i] = xyz[i * 3 + 0];
    py[i] = xyz[i * 3 + 1];
    pz[i] = xyz[i * 3 + 2];
    se[i] = std::sqrt(eps[i]);
  }
  for (int k = 0; k < 8; ++k) radii[k] = kProbeRadii[k];
  const float len = (float)(n - 1) * h;
  for (int k = 0; k < 3; ++k) {
    mid3[k] = box / 2;
    len3[k] = len;
  }
  const double gmin = (double)(mid3[0] - len3[0] / 2);
  const double rc_lj = 8.0, rc_el = 12.0;
  // Cell list with cell edge rc_el / 2.
  const double cell = rc_el / 2;
  const int nc = std::max(1, (int)std::ceil(box / cell));
  std::vector<std::vector<int>> cells((size_t)nc * nc * nc);
  auto cidx = [&](double v) { return std::min(nc - 1, std::max(0, (int)std::floor(v / cell))); };
  for (int i = 0; i < natom; ++i) cells[((size_t)cidx(px[i]) * nc + cidx(py[i])) * nc + cidx(pz[i])].push_back(i);
  const size_t G = (size_t)n * n * n;
  auto work = [&](int x0, int x1) {
    for (int ix = x0; ix < x1; ++ix)
      for (int iy = 0; iy < n; ++iy)
        for (int iz = 0; iz < n; ++iz) {
          const double gx = gmin + ix * (double)h, gy = gmin + iy * (double)h, gz = gmin + iz * (double)h;
          double lj[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          double el = 0;
          const int cx = cidx(gx), cy = cidx(gy), cz = cidx(gz);
          for (int ax = std::max(0, cx - 2); ax <= std::min(nc - 1, cx + 2); ++ax)
            for (int ay = std::max(0, cy - 2); ay <= std::min(nc - 1, cy + 2); ++ay)
              for (int az = std::max(0, cz - 2); az <= std::min(nc - 1, cz + 2); ++az)
                for (int i : cells[((size_t)ax * nc + ay) * nc + az]) {
                  const double dx = px[i] - gx, dy = py[i] - gy, dz = pz[i] - gz;
                  double d2 = dx * dx + dy * dy + dz * dz;
                  if (d2 > rc_el * rc_el) continue;
                  d2 = std::max(d2, 1e-4);
                  el += 332.0636 * q[i] / (4.0 * d2);
                  if (d2 > rc_lj * rc_lj) continue;
                  const double inv2 = 1.0 / d2;
                  for (int p = 0; p < 8; ++p) {
                    const double R = r[i] + kProbeRadii[p];
                    const double s2 = R * R * inv2;
                    const double s6 = s2 * s2 * s2;
                    lj[p] += se[i] * (s6 * s6 - 2.0 * s6);
                  }
                }
          const size_t gidx = ((size_t)ix * n + iy) * n + iz;
          for (int p = 0; p < 8; ++p) values[(size_t)p * G + gidx] = (float)std::min(lj[p], 3.0);
          values[(size_t)8 * G + gidx] = (float)std::max(-20.0, std::min(el, 20.0));
        }
  };
  const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) th.emplace_back(work, (int)((long)n * t / nt), (int)((long)n * (t + 1) / nt));
  for (auto& t : th) t.join();
}

static void centre(float* xyz, int natom) {
  double c[3] = {0, 0, 0};
  for (int a = 0; a < natom; ++a)
    for (int k = 0; k < 3; ++k) c[k] += xyz[a * 3 + k];
  for (int k = 0; k < 3; ++k) c[k] /= natom;
  for (int a = 0; a < natom; ++a)
    for (int k = 0; k < 3; ++k) xyz[a * 3 + k] = (float)(xyz[a * 3 + k] - c[k]);
}

static bool clash(const std::vector<double>& p, int natom, double dmin) {
  for (int i = 0; i < natom; ++i)
    for (int j = i + 2; j < natom; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) {
        const double d = p[i * 3 + k] - p[j * 3 + k];
        s += d * d;
      }
      if (s < dmin * dmin) return true;
    }
  return false;
}

void synth_ligand(uint64_t seed, int natom, float* coords, float* charge, float* eps, float* radius) {
  Xoshiro rng(seed);
  std::vector<double> p((size_t)natom * 3, 0.0);
  for (int i = 1; i < natom; ++i) {
    for (int attempt = 0;; ++attempt) {
      double v[3] = {rng.normal(), rng.normal(), rng.normal()};
      const double nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      if (nv < 1e-12) continue;
      double c[3];
      for (int k = 0; k < 3; ++k) c[k] = p[(i - 1) * 3 + k] + 1.5 * v[k] / nv;
      bool ok = true;
      for (int j = 0; j + 1 < i && ok; ++j) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += (c[k] - p[j * 3 + k]) * (c[k] - p[j * 3 + k]);
        ok = s >= 1.2 * 1.2;
      }
      if (ok || attempt > 10000) {
        for (int k = 0; k < 3; ++k) p[i * 3 + k] = c[k];
        break;
      }
    }
  }
  for (int i = 0; i < natom * 3; ++i) coords[i] = (float)p[i];
  centre(coords, natom);
  double qs = 0;
  for (int i = 0; i < natom; ++i) {
    charge[i] = (float)(0.25 * rng.normal());
    qs += charge[i];
  }
  for (int i = 0; i < natom; ++i) charge[i] = (float)(charge[i] - qs / natom);
  for (int i = 0; i < natom; ++i) {
    radius[i] = kProbeRadii[(int)(rng.uniform() * 8) & 7];
    eps[i] = (float)(0.05 + 0.15 * rng.uniform());
  }
}

// Rotation of atoms after bond (b, b+1) about that bond's axis by angle t.
static void torsion(std::vector<double>& p, int natom, int b, double t) {
  const double* a0 = &p[b * 3];
  const double* a1 = &p[(b + 1) * 3];
  double u[3] = {a1[0] - a0[0], a1[1] - a0[1], a1[2] - a0[2]};
  const double nu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int k = 0; k < 3; ++k) u[k] /= nu;
  const double c = std::cos(t), s = std::sin(t);
  const double o[3] = {a1[0], a1[1], a1[2]};
  for (int i = b + 2; i < natom; ++i) {
    double v[3] = {p[i * 3] - o[0], p[i * 3 + 1] - o[1], p[i * 3 + 2] - o[2]};
    const double dot = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
    const double cr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    for (int k = 0; k < 3; ++k) p[i * 3 + k] = o[k] + v[k] * c + cr[k] * s + u[k] * dot * (1 - c);
  }
}

void synth_conformers(uint64_t seed, int natom, const float* coords, int nconf, float* out) {
  Xoshiro rng(seed);
  for (int c = 0; c < nconf; ++c) {
    float* o = out + (size_t)c * natom * 3;
    std::vector<double> p(coords, coords + (size_t)natom * 3);
    if (c > 0) {
      for (int attempt = 0; attempt < 200; ++attempt) {
        std::vector<double> t(p);
        for (int b = 1; b + 2 < natom; ++b) torsion(t, natom, b, 0.6 * rng.normal());
        if (!clash(t, natom, 1.2)) {
          p = t;
          break;
        }
      }
    }
    for (int i = 0; i < natom * 3; ++i) o[i] = (float)p[i];
    centre(o, natom);
  }
}

// ---------------------------------------------------------------------------
// Grid text format, ReadGrids.cpp:21-66: 3 title lines; numGrids X Y Z
// midx midy midz xlen ylen zlen spacing restraint_k; numGrids-1 radii; per
// grid one skipped line then "x y z value" lines, Z fastest.
// ---------------------------------------------------------------------------
void read_grid_header(const std::string& path, GridHeader& g) {
  std::ifstream in(path);
  if (!in) throw Error(kErrInvalid, "cannot open grid file " + path);
  for (int i = 0; i < 3; ++i) in.ignore(1 << 20, '\n');
  in >> g.num_grids >> g.dims[0] >> g.dims[1] >> g.dims[2] >> g.mid[0] >> g.mid[1] >> g.mid[2] >> g.len[0] >>
      g.len[1] >> g.len[2] >> g.spacing >> g.restraint_k;
  if (!in || g.num_grids < 2 || g.dims[0] <= 0 || g.dims[1] <= 0 || g.dims[2] <= 0)
    throw Error(kErrInvalid, "malformed grid header in " + path);
  g.radii.resize(g.num_grids - 1);
  for (auto& r : g.radii) in >> r;
  if (!in) throw Error(kErrInvalid, "malformed grid radii in " + path);
}

void read_grids(const std::string& path, GridHeader& g, std::vector<float>& values) {
  read_grid_header(path, g);
  std::ifstream in(path);
  for (int i = 0; i < 3; ++i) in.ignore(1 << 20, '\n');
  GridHeader tmp;
  in >> tmp.num_grids >> tmp.dims[0] >> tmp.dims[1] >> tmp.dims[2] >> tmp.mid[0] >> tmp.mid[1] >> tmp.mid[2] >>
      tmp.len[0] >> tmp.len[1] >> tmp.len[2] >> tmp.spacing >> tmp.restraint_k;
  for (int i = 0; i < g.num_grids - 1; ++i) {
    float r;
    in >> r;
  }
  in.ignore(1 << 20, '\n');
  const size_t G = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
  values.resize(G * g.num_grids);
  std::string line;
  size_t n = 0;
  for (int k = 0; k < g.num_grids; ++k) {
    std::getline(in, line);  // per-grid separator line
    for (size_t j = 0; j < G; ++j) {
      if (!std::getline(in, line)) throw Error(kErrInvalid, "grid file truncated: " + path);
      float x, y, z, v;
      if (std::sscanf(line.c_str(), "%f %f %f %f", &x, &y, &z, &v) != 4)
        throw Error(kErrInvalid, "malformed grid value line in " + path);
      values[n++] = v;
    }
  }
}

void write_grids(const std::string& path, const GridHeader& g, const float* values) {
  std::FILE* f = std::fopen(path.c_str(), "w");
  if (!f) throw Error(kErrInvalid, "cannot write grid file " + path);
  std::fprintf(f, "* potential grids\n* written by cdk\n*\n");
  std::fprintf(f, "%d %d %d %d %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", g.num_grids, g.dims[0], g.dims[1],
               g.dims[2], g.mid[0], g.mid[1], g.mid[2], g.len[0], g.len[1], g.len[2], g.spacing, g.restraint_k);
  for (int i = 0; i < g.num_grids - 1; ++i) std::fprintf(f, "%.9g ", g.radii[i]);
  std::fprintf(f, "\n");
  const size_t G = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
  const double x0 = (double)(g.mid[0] - g.len[0] / 2), y0 = (double)(g.mid[1] - g.len[1] / 2),
               z0 = (double)(g.mid[2] - g.len[2] / 2);
  for (int k = 0; k < g.num_grids; ++k) {
    std::fprintf(f, "grid %d\n", k);
    size_t j = 0;
    for (int ix = 0; ix < g.dims[0]; ++ix)
      for (int iy = 0; iy < g.dims[1]; ++iy)
        for (int iz = 0; iz < g.dims[2]; ++iz, ++j)
          std::fprintf(f, "%.6f %.6f %.6f %.9g\n", x0 + ix * g.spacing, y0 + iy * g.spacing, z0 + iz * g.spacing,
                       values[(size_t)k * G + j]);
  }
  std::fclose(f);
}

}  // namespace cdk
