// Irregular graphs from point sets: seeded Poisson-disk sampling in a
// periodic square box and neighbour tables by distance shells.
//
// The reference reads such graphs as tabulated tables (its quasicrystal
// input, nborsList_AppxtSz(N).bin with fixed widths 8/3/8,
// CF/subrouts.cpp:1012); here the widths are the largest shell counts found,
// and the result is an ordinary tabulated lattice (lattice_from_tables).
#include <math.h>

#include <algorithm>
#include <vector>

#include "hsmd_internal.h"

namespace hsmd {

namespace {

// splitmix64 (Steele, Lea and Flood, OOPSLA 2014): a seeded host stream.
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

double wrap(double x, double box) {
  x = fmod(x, box);
  return x < 0.0 ? x + box : (x >= box ? x - box : x);
}

double min_image(double d, double box) {
  d = fmod(d, box);
  if (d > 0.5 * box) d -= box;
  else if (d < -0.5 * box) d += box;
  return d;
}

}  // namespace

// Bridson, "Fast Poisson disk sampling in arbitrary dimensions" (SIGGRAPH
// 2007 sketches): a background grid of cell side <= r/sqrt(2) holds at most
// one point; an active point spawns up to k = 30 candidates in the annulus
// [r, 2r); a candidate is kept when no point lies within r (periodic
// minimum-image distance).  Returns the number of points written (<= max_n).
int poisson_disk(double box, double rmin, uint64_t seed, int max_n, double *out) {
  if (!(box > 0.0) || !(rmin > 0.0) || rmin * 4.0 > box)
    throw Error(HSMD_ERR_ARG, "Poisson disk: need box > 4 r_min > 0");
  if (max_n < 1) throw Error(HSMD_ERR_ARG, "Poisson disk: max_points must be >= 1");
  const int ng = std::max(1, (int)floor(box / (rmin / sqrt(2.0))));
  const double cs = box / ng;
  std::vector<int> grid((size_t)ng * ng, -1);
  std::vector<double> pts;
  std::vector<int> active;
  SplitMix rng{seed ^ 0x5851f42d4c957f2dull};
  auto cell = [&](double x) { return std::min(ng - 1, (int)(x / cs)); };
  auto add = [&](double x, double y) {
    const int id = (int)(pts.size() / 2);
    pts.push_back(x);
    pts.push_back(y);
    grid[(size_t)cell(y) * ng + cell(x)] = id;
    active.push_back(id);
  };
  auto free_at = [&](double x, double y) {
    const int cx = cell(x), cy = cell(y);
    for (int dy = -2; dy <= 2; dy++)
      for (int dx = -2; dx <= 2; dx++) {
        const int gx = ((cx + dx) % ng + ng) % ng, gy = ((cy + dy) % ng + ng) % ng;
        const int id = grid[(size_t)gy * ng + gx];
        if (id < 0) continue;
        const double ddx = min_image(pts[2 * id] - x, box), ddy = min_image(pts[2 * id + 1] - y, box);
        if (ddx * ddx + ddy * ddy < rmin * rmin) return false;
      }
    return true;
  };
  add(rng.uniform() * box, rng.uniform() * box);
  const double tp = 2.0 * acos(-1.0);
  while (!active.empty() && (int)(pts.size() / 2) < max_n) {
    const size_t a = (size_t)(rng.uniform() * active.size()) % active.size();
    const int id = active[a];
    bool found = false;
    for (int k = 0; k < 30 && !found; k++) {
      const double rad = rmin * (1.0 + rng.uniform()), ang = tp * rng.uniform();
      const double x = wrap(pts[2 * id] + rad * cos(ang), box);
      const double y = wrap(pts[2 * id + 1] + rad * sin(ang), box);
      if (free_at(x, y)) {
        add(x, y);
        found = true;
      }
    }
    if (!found) {
      active[a] = active.back();
      active.pop_back();
    }
  }
  const int n = (int)(pts.size() / 2);
  std::copy(pts.begin(), pts.end(), out);
  return n;
}

// Neighbour tables of points in a periodic box: shell 1 holds the points at
// minimum-image distance r < r1, shell 2 r1 <= r < r2, shell X r2 <= r < r3;
// inside a shell, slots go by distance then index.  Table widths are the
// largest counts; missing slots are -1.
Lattice *lattice_from_points(int n, const double *rvec, double box, const double r[3],
                             const hsmd_couplings &c) {
  if (n < 2 || !rvec) throw Error(HSMD_ERR_ARG, "need at least two points");
  if (!(r[0] > 0.0 && r[1] >= r[0] && r[2] >= r[1])) throw Error(HSMD_ERR_ARG, "shell radii must be increasing");
  if (!(box > 2.0 * r[2])) throw Error(HSMD_ERR_ARG, "box must exceed twice the outer shell radius");
  const int ng = std::max(1, (int)floor(box / r[2]));
  const double cs = box / ng;
  std::vector<std::vector<int>> cells((size_t)ng * ng);
  auto cell = [&](double x) { return std::min(ng - 1, std::max(0, (int)(wrap(x, box) / cs))); };
  for (int i = 0; i < n; i++) cells[(size_t)cell(rvec[2 * i + 1]) * ng + cell(rvec[2 * i])].push_back(i);
  std::vector<std::vector<std::pair<double, int>>> sh[3];
  for (int s = 0; s < 3; s++) sh[s].resize(n);
  const int reach = ng >= 3 ? 1 : 0;
  for (int i = 0; i < n; i++) {
    const int cx = cell(rvec[2 * i]), cy = cell(rvec[2 * i + 1]);
    std::vector<int> seen;
    for (int dy = -reach; dy <= reach; dy++)
      for (int dx = -reach; dx <= reach; dx++) {
        const int gx = ((cx + dx) % ng + ng) % ng, gy = ((cy + dy) % ng + ng) % ng;
        for (int j : cells[(size_t)gy * ng + gx]) seen.push_back(j);
      }
    if (reach == 0) { seen.clear(); for (int j = 0; j < n; j++) seen.push_back(j); }
    std::sort(seen.begin(), seen.end());
    seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
    for (int j : seen) {
      if (j == i) continue;
      const double dx = min_image(rvec[2 * j] - rvec[2 * i], box);
      const double dy = min_image(rvec[2 * j + 1] - rvec[2 * i + 1], box);
      const double d = sqrt(dx * dx + dy * dy);
      const int s = d < r[0] ? 0 : (d < r[1] ? 1 : (d < r[2] ? 2 : -1));
      if (s >= 0) sh[s][i].push_back({d, j});
    }
  }
  int zn[3] = {0, 0, 0};
  for (int s = 0; s < 3; s++)
    for (int i = 0; i < n; i++) {
      std::sort(sh[s][i].begin(), sh[s][i].end());
      zn[s] = std::max(zn[s], (int)sh[s][i].size());
    }
  const int Zt = zn[0] + zn[1] + zn[2];
  if (Zt == 0) throw Error(HSMD_ERR_ARG, "no neighbours within the shells");
  std::vector<int32_t> nb((size_t)n * Zt, -1);
  for (int i = 0; i < n; i++) {
    int off = 0;
    for (int s = 0; s < 3; s++) {
      for (size_t k = 0; k < sh[s][i].size(); k++) nb[(size_t)i * Zt + off + k] = sh[s][i][k].second;
      off += zn[s];
    }
  }
  return lattice_from_tables(n, zn[0], zn[1], zn[2], nb.data(), rvec, c);
}

}  // namespace hsmd
