// Host check of the workgroup-to-XCD maps the kz-plane and line-group
// kernels use (cdocker_amd/csrc/xcd_map.hpp): for every grid edge's item
// count (N/2 + 1 kz planes; the plane form's line groups) and a range of
// unit counts, each map's grid covers every (item, unit) pair exactly once
// and nothing out of range, and the plane-major map gives the 8 XCDs
// (workgroup w on XCD w % 8) work that differs by at most one item.  Prints
// "ok" and exits 0, or names the failing case and exits 1.
#include <cstdio>
#include <vector>

#include "xcd_map.hpp"

namespace {

int fail(const char* map, int M, int G, const char* what) {
  std::printf("FAIL %s M=%d G=%d: %s\n", map, M, G, what);
  return 1;
}

template <class Of>
int check(const char* map, int M, int G, unsigned grid, Of of, bool balanced) {
  std::vector<int> seen((size_t)M * G, 0);
  long per_xcd[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (unsigned w = 0; w < grid; ++w) {
    int item = -1, unit = -1;
    if (!of((int)w, item, unit)) continue;
    if (item < 0 || item >= M || unit < 0 || unit >= G) return fail(map, M, G, "out of range");
    if (seen[(size_t)item * G + unit]++) return fail(map, M, G, "pair covered twice");
    ++per_xcd[w & 7];
  }
  for (int v : seen)
    if (v != 1) return fail(map, M, G, "pair not covered");
  if (balanced) {
    long lo = per_xcd[0], hi = per_xcd[0];
    for (long c : per_xcd) {
      lo = c < lo ? c : lo;
      hi = c > hi ? c : hi;
    }
    if (hi - lo > 1) return fail(map, M, G, "XCD loads differ by more than one item");
  }
  return 0;
}

}  // namespace

int main() {
  std::vector<int> items;
  const int edges[] = {32, 48, 64, 96, 128};
  for (int N : edges) {
    items.push_back(N / 2 + 1);  // kz planes of the half spectrum
    const int L = (N == 64 || N == 128) ? 8 : 4;
    const int lpb = 4 * (64 / L);
    items.push_back((N * (N / 2 + 1) + lpb - 1) / lpb);  // the plane form's line groups
  }
  std::vector<int> units;
  for (int g = 1; g <= 80; ++g) units.push_back(g);
  const int big[] = {127, 128, 129, 255, 316, 342, 1081, 4097};
  for (int g : big) units.push_back(g);
  int bad = 0;
  for (int M : items) {
    const int Q = M / 8, R = M % 8;
    for (int G : units) {
      bad |= check("unit", M, G, cdk::xcd_unit_grid(Q, R, G),
                   [&](int w, int& it, int& u) { return cdk::xcd_unit_of(w, Q, R, G, it, u); }, false);
      bad |= check("plane", M, G, cdk::xcd_plane_grid(Q, R, G),
                   [&](int w, int& it, int& u) { return cdk::xcd_plane_of(w, Q, R, G, it, u); }, true);
      if (bad) return 1;
    }
  }
  std::printf("ok\n");
  return 0;
}
