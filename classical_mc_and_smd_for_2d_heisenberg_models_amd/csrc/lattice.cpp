// Host-side lattice construction for the hsmd library.
//
// Site numbering, neighbour shells and the in-shell slot order follow the
// reference exactly, because the kernels' summation order and the
// nborsList*.bin drop-in format depend on them:
//   site coordinates      make_rvecList   CF/subrouts.cpp:2714-2975
//   neighbour shells      make_nborsTable CF/subrouts.cpp:2983-3316
//   bond list             make_bondList   CF/subrouts.cpp:3426-3560
//   coupling flags        convert_inputs2data CF/subrouts.cpp:1087-2136
// The neighbour search is the reference's integer-coordinate criterion
// (spacing scaled by 500, dist = floor(|dr|), shells split at 550 / 742 or 909
// / 1050) evaluated over a cell list instead of the reference's O(N^2 * 9)
// scan; candidates are visited in the reference's extended-list order so the
// kagome angle sums round identically.
#include <math.h>

#include <algorithm>
#include <map>
#include <unordered_map>

#include "hsmd_internal.h"

namespace hsmd {

static const double kPi = acos(-1.0);
static const double kSq2 = sqrt(2.0);
static const double kSq3 = sqrt(3.0);
static const double kC60 = cos(kPi / 3.0);
static const double kS60 = sin(kPi / 3.0);

static bool not_zero(double v) { return fabs(v) > 1.0e-12; }

void derive_flags(Lattice &lat) {
  const hsmd_couplings &c = lat.c;
  lat.disorder = c.imp_ratio > 0.0;
  if (c.imp_ratio < 0.0 || c.imp_ratio >= 1.0)
    throw Error(HSMD_ERR_GUARD, "impurity ratio must be in [0,1)");
  lat.j2_on = not_zero(c.J2);
  lat.jx_on = not_zero(c.JX);
  lat.j2ab_on = false;
  if (not_zero(c.J2_delta)) {
    if (lat.geom != "lieb")
      throw Error(HSMD_ERR_GUARD, "J2_delta requires the lieb geometry");
    lat.j2_on = true;
    lat.j2ab_on = true;
  }
  if (lat.disorder) {
    if (not_zero(c.J2ic) || not_zero(c.J2ii)) lat.j2_on = true;
    if (not_zero(c.JXic) || not_zero(c.JXii)) lat.jx_on = true;
  }
  lat.xxz_on = (c.lambda1 != 1.0) || (c.lambda2 != 1.0);
  lat.h_on = not_zero(c.h);
  lat.c3sym = (!lat.tabulated) && lat.geom != "square" && lat.geom != "lieb";
}

double Lattice::bondJ(int shell, int i, int j) const {
  double J;
  const int a1 = imp.empty() ? 0 : imp[i], a2 = imp.empty() ? 0 : imp[j];
  if (shell == 1) {
    J = disorder ? (a1 == a2 ? (a1 ? c.J1ii : c.J1) : c.J1ic) : c.J1;
  } else if (shell == 2) {
    J = disorder ? (a1 == a2 ? (a1 ? c.J2ii : c.J2) : c.J2ic) : c.J2;
    if (j2ab_on) J += ((i & 1) ? -1 : 1) * ((j & 1) ? -1 : 1) * c.J2_delta;
  } else {
    J = disorder ? (a1 == a2 ? (a1 ? c.JXii : c.JX) : c.JXic) : c.JX;
  }
  return J;
}

// get_spinStiff (CF/subrouts.cpp:6594-6604): the shell coupling, or the
// disorder matrix entry, without the Lieb J2 asymmetry.
double Lattice::stiffJ(int shell, int i, int j) const {
  const int a1 = imp.empty() ? 0 : imp[i], a2 = imp.empty() ? 0 : imp[j];
  if (shell == 1) return disorder ? (a1 == a2 ? (a1 ? c.J1ii : c.J1) : c.J1ic) : c.J1;
  if (shell == 2) return disorder ? (a1 == a2 ? (a1 ? c.J2ii : c.J2) : c.J2ic) : c.J2;
  return disorder ? (a1 == a2 ? (a1 ? c.JXii : c.JX) : c.JXic) : c.JX;
}

// dq = 4 (2 pi) / Gsz for tabulated graphs, 2 pi / Gsz otherwise.
double Lattice::sq_dq() const {
  const double pi2 = 2.0 * acos(-1.0);
  return tabulated ? 4.0 * pi2 / Gsz : pi2 / Gsz;
}

double Lattice::slotJ(int shell, int k, int i, int nb) const {
  double J;
  const int a1 = imp.empty() ? 0 : imp[i], a2 = imp.empty() ? 0 : imp[nb];
  if (shell == 1) {
    J = disorder ? (a1 == a2 ? (a1 ? c.J1ii : c.J1) : c.J1ic) : c.J1;
  } else if (shell == 2) {
    J = disorder ? (a1 == a2 ? (a1 ? c.J2ii : c.J2) : c.J2ic) : c.J2;
    // Lieb asymmetry by site parity and slot parity (CF/subrouts.cpp:4527-4531)
    if (j2ab_on) J += ((i & 1) ? -1 : 1) * ((k & 1) ? -1 : 1) * c.J2_delta;
  } else {
    J = disorder ? (a1 == a2 ? (a1 ? c.JXii : c.JX) : c.JXic) : c.JX;
  }
  return J;
}

// ---------------------------------------------------------------- geometry

struct GeomInfo {
  int nsubs, gridFac, zn1, zn2, znx;
  double szFac;
};

static GeomInfo geom_info(const std::string &g) {
  if (g == "square") return {1, 1, 4, 4, 4, 1.0};
  if (g == "triang") return {1, 1, 6, 6, 6, 1.0};
  if (g == "hexagn") return {2, 3, 3, 6, 3, kSq3};
  if (g == "kagome") return {3, 2, 4, 4, 6, 2.0};
  if (g == "lieb") return {2, 2, 4, 4, 4, kSq2};
  throw Error(HSMD_ERR_ARG, "unknown geometry: " + g);
}

static void make_sites(Lattice &lat, const GeomInfo &gi) {
  const int L = lat.L;
  const double sz = gi.szFac;
  lat.rvec.assign((size_t)lat.N * 2, 0.0);
  int k = 0;
  for (int i = 0; i < L; i++) {
    for (int j = 0; j < L; j++) {
      double r0, r1;
      if (lat.geom == "square" || lat.geom == "lieb") {
        r0 = i * sz;
        r1 = j * sz;
      } else {
        r0 = sz * (i * 1.0 + j * kC60);
        r1 = sz * (i * 0.0 + j * kS60);
      }
      auto put = [&](double x, double y) {
        lat.rvec[2 * k] = x;
        lat.rvec[2 * k + 1] = y;
        k++;
      };
      if (lat.geom == "square" || lat.geom == "triang") {
        put(r0, r1);
      } else if (lat.geom == "hexagn") {
        put(r0 + 0.5 * kSq3, r1 + 0.5 * 1.0);
        put(r0 + 1.0 * kSq3, r1 + 1.0 * 1.0);
      } else if (lat.geom == "kagome") {
        for (int n = 0; n < 3; n++) {
          const double a = kPi - n * (kPi / 3.0);
          put(r0 + cos(a), r1 + sin(a));
        }
      } else {  // lieb
        for (int n = 0; n < 2; n++) {
          const double a = kPi - n * (kPi / 2.0);
          put(r0 + 0.5 * sz * cos(a), r1 + 0.5 * sz * sin(a));
        }
      }
    }
  }
  if (k != lat.N) throw Error(HSMD_ERR_ARG, "site count mismatch");
}

// DFT grid and its map to sites (CF/subrouts.cpp:2907-2975).
static void make_grid_map(Lattice &lat, const GeomInfo &gi) {
  const int G = lat.Gsz;
  const double fc = gi.szFac / gi.gridFac;
  int i0 = 0;
  const bool sq = (lat.geom == "square" || lat.geom == "lieb");
  if (lat.geom == "lieb" || lat.geom == "kagome") i0 = 1;
  // hash of site positions on a 1e-6 grid
  std::unordered_map<int64_t, int> pos;
  auto key = [](double x, double y) {
    const int64_t ix = llround(x * 1.0e6), iy = llround(y * 1.0e6);
    return (ix << 32) ^ (iy & 0xffffffffLL);
  };
  for (int k = 0; k < lat.N; k++) pos[key(lat.rvec[2 * k], lat.rvec[2 * k + 1])] = k;
  lat.gridMap.assign((size_t)G * G, -1);
  int cnt = 0;
  for (int i = 0; i < G; i++) {
    for (int j = 0; j < G; j++) {
      double x, y;
      if (sq) {
        x = (i - i0) * fc;
        y = j * fc;
      } else {
        x = fc * ((i - i0) * 1.0 + j * kC60);
        y = fc * ((i - i0) * 0.0 + j * kS60);
      }
      const int n = j + i * G;
      // probe the rounding neighbourhood, accept within 1e-8 as the reference
      for (int dx = -1; dx <= 1 && lat.gridMap[n] < 0; dx++)
        for (int dy = -1; dy <= 1 && lat.gridMap[n] < 0; dy++) {
          auto it = pos.find(key(x + dx * 1.0e-6, y + dy * 1.0e-6));
          if (it == pos.end()) continue;
          const int k = it->second;
          if (fabs(x - lat.rvec[2 * k]) < 1.0e-8 && fabs(y - lat.rvec[2 * k + 1]) < 1.0e-8) {
            lat.gridMap[n] = k;
            cnt++;
          }
        }
    }
  }
  if (cnt != lat.N) throw Error(HSMD_ERR_ARG, "DFT grid does not cover every site");
}

struct Cand {
  int m;        // index in the reference's extended list
  int tag;      // site number (or -1 ghost)
  double angle;
  int shell;
};

static void make_nbors(Lattice &lat, const GeomInfo &gi, bool pbc) {
  const double spc = 500.0;
  const int N = lat.N, L = lat.L;
  const double sz = gi.szFac;
  const bool sq = (lat.geom == "square" || lat.geom == "lieb");
  // integer site coordinates
  std::vector<int> px(N), py(N);
  for (int k = 0; k < N; k++) {
    px[k] = (int)lround(spc * lat.rvec[2 * k]);
    py[k] = (int)lround(spc * lat.rvec[2 * k + 1]);
  }
  int nx1, nx2, ny;
  if (sq) {
    nx1 = nx2 = ny = (int)lround(L * sz * spc);
  } else {
    nx1 = (int)lround(L * sz * (spc * 1.0));
    nx2 = (int)lround(L * sz * (spc * kC60));
    ny = (int)lround(L * sz * (spc * 0.0 + spc * kS60)) - 1;
  }
  int sh[8][2] = {{+nx1, 0}, {-nx1, 0}, {nx2, +ny}, {nx2, -ny},
                  {-nx2, +ny}, {-nx2, -ny}, {0, +ny}, {0, -ny}};
  if (!sq) {
    sh[6][0] = nx2 + nx1; sh[6][1] = ny;
    sh[7][0] = -nx2 - nx1; sh[7][1] = -ny;
  }
  // extended list: originals (m < N), then for each site n its 8 images
  const int M = 9 * N;
  std::vector<int> ex(M), ey(M), ez(M);
  for (int k = 0; k < N; k++) { ex[k] = px[k]; ey[k] = py[k]; ez[k] = k; }
  for (int n = 0; n < N; n++)
    for (int i = 0; i < 8; i++) {
      const int m = N + 8 * n + i;
      ex[m] = sh[i][0] + px[n];
      ey[m] = sh[i][1] + py[n];
      ez[m] = n;
    }
  const int rA1 = (int)floor(spc * 1.10);
  const int rA3 = (int)floor(2.0 * spc * 1.05);
  const int rA2 = sq ? (int)floor(kSq2 * spc * 1.05) : (int)floor(kSq3 * spc * 1.05);
  // cell list over the extended points
  const int cs = rA3 + 1;
  int xmin = ex[0], ymin = ey[0], xmax = ex[0], ymax = ey[0];
  for (int m = 0; m < M; m++) {
    xmin = std::min(xmin, ex[m]); xmax = std::max(xmax, ex[m]);
    ymin = std::min(ymin, ey[m]); ymax = std::max(ymax, ey[m]);
  }
  const int ncx = (xmax - xmin) / cs + 1, ncy = (ymax - ymin) / cs + 1;
  std::vector<int> head((size_t)ncx * ncy + 1, 0), list(M);
  auto cell = [&](int x, int y) { return ((x - xmin) / cs) * ncy + (y - ymin) / cs; };
  for (int m = 0; m < M; m++) head[cell(ex[m], ey[m]) + 1]++;
  for (size_t c = 1; c < head.size(); c++) head[c] += head[c - 1];
  {
    std::vector<int> fill(head.begin(), head.end() - 1);
    for (int m = 0; m < M; m++) list[fill[cell(ex[m], ey[m])]++] = m;
  }
  const int Zt = lat.zn1 + lat.zn2 + lat.znx;
  lat.nbors.assign((size_t)N * Zt, -1);
  std::vector<Cand> cands;
  for (int n = 0; n < N; n++) {
    const int x0 = ex[n], y0 = ey[n];
    cands.clear();
    const int cx = (x0 - xmin) / cs, cy = (y0 - ymin) / cs;
    for (int a = cx - 1; a <= cx + 1; a++) {
      if (a < 0 || a >= ncx) continue;
      for (int b = cy - 1; b <= cy + 1; b++) {
        if (b < 0 || b >= ncy) continue;
        const int c = a * ncy + b;
        for (int p = head[c]; p < head[c + 1]; p++) {
          const int m = list[p];
          if (ez[m] == n) continue;
          const double ddx = (double)(ex[m] - x0), ddy = (double)(ey[m] - y0);
          const int dist = (int)floor(sqrt(pow(ddx, 2) + pow(ddy, 2)));
          int shell = 0;
          if (dist < rA1) shell = 1;
          else if (dist > rA1 && dist < rA2) shell = 2;
          else if (dist > rA2 && dist < rA3) shell = 3;
          if (!shell) continue;
          int tag = ez[m];
          if (m >= N && !pbc) tag = -1;
          cands.push_back({m, tag, atan2(ddy, ddx), shell});
        }
      }
    }
    std::sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.m < b.m; });
    std::vector<std::pair<int, double>> s1, s2, s3;  // (tag, angle)
    double angleSum = 0.0;
    for (const Cand &c : cands) {
      if (c.shell == 1) { angleSum += c.angle; s1.push_back({c.tag, c.angle}); }
      else if (c.shell == 2) s2.push_back({c.tag, c.angle});
      else s3.push_back({c.tag, c.angle});
    }
    if ((int)s1.size() != lat.zn1 || (int)s2.size() != lat.zn2 || (int)s3.size() != lat.znx)
      throw Error(HSMD_ERR_ARG, "neighbour count mismatch at site " + std::to_string(n) +
                                    " (lattice too small?)");
    auto desc = [](const std::pair<int, double> &a, const std::pair<int, double> &b) {
      return a.second > b.second;
    };
    std::sort(s1.begin(), s1.end(), desc);
    std::sort(s2.begin(), s2.end(), desc);
    std::sort(s3.begin(), s3.end(), desc);
    if (lat.geom == "hexagn") {
      std::rotate(s2.begin(), s2.begin() + ((n % 2 == 0) ? 1 : 2), s2.end());
    } else if (lat.geom == "kagome") {
      int ic1 = 1, ic2 = 1;
      const double avg = angleSum * (1.0 / (double)s1.size());
      if (avg > 0.0) {
        if (avg > 1.0) ic2 -= 1;
        else { ic1 -= 1; ic2 -= 1; }
      }
      std::rotate(s1.begin(), s1.begin() + ic1, s1.end());
      std::rotate(s2.begin(), s2.begin() + ic2, s2.end());
    }
    int32_t *row = &lat.nbors[(size_t)n * Zt];
    for (int k = 0; k < lat.zn1; k++) row[k] = s1[k].first;
    for (int k = 0; k < lat.zn2; k++) row[lat.zn1 + k] = s2[k].first;
    for (int k = 0; k < lat.znx; k++) row[lat.zn1 + lat.zn2 + k] = s3[k].first;
  }
}

// Order-parameter labellings (CF/subrouts.cpp:5382-5800).
static void make_labels(Lattice &lat) {
  Labels &lb = lat.labels;
  const int N = lat.N, L = lat.L;
  lb.present = true;
  lb.stag.assign(N, -1); lb.hs.assign(N, -1); lb.vs.assign(N, -1); lb.ds.assign(N, -1);
  lb.n_stag = lat.c3sym ? 3 : 2;
  const int ifc = lat.c3sym ? 1 : 0;
  if (lat.geom != "lieb") {
    for (int i = 0; i < L; i++)
      for (int j = 0; j < L; j++) {
        const int st = i * L + j;
        lb.stag[st] = (int8_t)((i + j + ifc * i) % lb.n_stag);
        lb.hs[st] = (int8_t)(j % 2);
        lb.vs[st] = (int8_t)(i % 2);
        lb.ds[st] = (int8_t)((i + j) % 2);
      }
  } else {
    int n = 0;
    for (int i = 0; i < L; i++)
      for (int j = 0; j < L; j++) {
        lb.hs[n] = lb.hs[n + 1] = (int8_t)((i + j) & 1);
        lb.vs[n] = (int8_t)((i + j) % 2);
        lb.vs[n + 1] = (int8_t)(1 - (i + j) % 2);
        n += 2;
      }
    for (int k = 0; k < N; k++) { lb.stag[k] = (int8_t)(k % 2); lb.ds[k] = (int8_t)(k % 2); }
  }
  lb.c3 = lat.c3sym;
  lb.z2 = (lat.geom == "square");
  if (lb.z2) {
    const int Zt = lat.zn1 + lat.zn2 + lat.znx;
    lb.z2nb.assign((size_t)3 * N, 0);
    for (int i = 0; i < N; i++) {
      lb.z2nb[i] = lat.nbors[(size_t)i * Zt + 0];
      lb.z2nb[N + i] = lat.nbors[(size_t)i * Zt + 1];
      lb.z2nb[2 * N + i] = lat.nbors[(size_t)i * Zt + lat.zn1 + 0];
    }
  }
}

Lattice *lattice_generate(const std::string &geom, int L, bool pbc, bool strict,
                          const hsmd_couplings &c) {
  GeomInfo gi = geom_info(geom);
  if (L < 2) throw Error(HSMD_ERR_ARG, "L must be >= 2");
  Lattice *lat = new Lattice();
  try {
    lat->geom = geom;
    lat->L = L;
    lat->c = c;
    lat->nsubs = gi.nsubs;
    lat->gridFac = gi.gridFac;
    lat->zn1 = gi.zn1; lat->zn2 = gi.zn2; lat->znx = gi.znx;
    lat->N = gi.nsubs * L * L;
    lat->Gsz = gi.gridFac * L;
    lat->Nq = lat->Gsz * lat->Gsz;
    derive_flags(*lat);
    if (strict) {
      if (L % 2 != 0) throw Error(HSMD_ERR_GUARD, "Lattice length must multiple of 2");
      if (L % 3 != 0 && lat->c3sym) throw Error(HSMD_ERR_GUARD, "Lattice length must multiple of 3");
    }
    make_sites(*lat, gi);
    make_grid_map(*lat, gi);
    make_nbors(*lat, gi, pbc);
    lat->pbc = pbc;
    lat->imp.assign(lat->N, 0);
    lat->vacant.assign(lat->N, 0);
    make_labels(*lat);
    lat->rebuild();
  } catch (...) {
    delete lat;
    throw;
  }
  return lat;
}

Lattice *lattice_from_tables(int n, int zn1, int zn2, int znx, const int32_t *nb,
                             const double *rvec, const hsmd_couplings &c) {
  if (n <= 0 || zn1 < 0 || zn2 < 0 || znx < 0 || !nb)
    throw Error(HSMD_ERR_ARG, "invalid table shape");
  Lattice *lat = new Lattice();
  try {
    lat->geom = "tables";
    lat->tabulated = true;
    lat->c = c;
    lat->N = n;
    lat->zn1 = zn1; lat->zn2 = zn2; lat->znx = znx;
    // qcrystal grid side: even integer part of sqrt(N) (CF/subrouts.cpp:1178-1182)
    lat->L = ((int)sqrt((double)n) / 2) * 2;
    lat->Gsz = lat->L;
    lat->Nq = lat->L * lat->L;
    const int Zt = zn1 + zn2 + znx;
    lat->nbors.assign(nb, nb + (size_t)n * Zt);
    for (int32_t v : lat->nbors)
      if (v >= n) throw Error(HSMD_ERR_ARG, "neighbour index out of range");
    if (rvec) lat->rvec.assign(rvec, rvec + (size_t)2 * n);
    derive_flags(*lat);
    lat->imp.assign(n, 0);
    lat->vacant.assign(n, 0);
    lat->rebuild();
  } catch (...) {
    delete lat;
    throw;
  }
  return lat;
}

// ------------------------------------------------------------- derived data

void Lattice::rebuild() {
  const int Zt = zn1 + zn2 + znx;
  if (!disorder) {
    c.J1ic = c.J1ii = c.J2ic = c.J2ii = c.JXic = c.JXii = 0.0;
  }
  // effective table with vacancies removed
  std::vector<int32_t> tab = nbors;
  bool any_vac = false;
  for (int i = 0; i < N; i++) any_vac |= (vacant[i] != 0);
  if (any_vac) {
    for (int i = 0; i < N; i++)
      for (int k = 0; k < Zt; k++) {
        int32_t &v = tab[(size_t)i * Zt + k];
        if (vacant[i] || (v >= 0 && vacant[v])) v = -1;
      }
  }
  // bond list, reference order: shell 1, then 2 (if on), then X (if on);
  // an unordered pair is kept once (CF/subrouts.cpp:3426-3560)
  bi.clear(); bj.clear(); bshell.clear();
  {
    std::vector<std::vector<int32_t>> seen(N);
    auto add_shell = [&](int shell, int off, int zs) {
      for (int i = 0; i < N; i++)
        for (int k = 0; k < zs; k++) {
          const int nb = tab[(size_t)i * Zt + off + k];
          if (nb < 0) continue;
          const int a = std::min(i, nb), b = std::max(i, nb);
          auto &v = seen[a];
          if (std::find(v.begin(), v.end(), b) != v.end()) continue;
          v.push_back(b);
          bi.push_back(i); bj.push_back(nb); bshell.push_back(shell);
        }
    };
    add_shell(1, 0, zn1);
    if (j2_on) add_shell(2, zn1, zn2);
    if (jx_on) add_shell(3, zn1 + zn2, znx);
  }
  // active slots
  slot_shell.clear(); slot_k.clear();
  for (int k = 0; k < zn1; k++) { slot_shell.push_back(1); slot_k.push_back(k); }
  if (j2_on) for (int k = 0; k < zn2; k++) { slot_shell.push_back(2); slot_k.push_back(k); }
  if (jx_on) for (int k = 0; k < znx; k++) { slot_shell.push_back(3); slot_k.push_back(k); }
  Z = (int)slot_shell.size();
  if (Z > ZMAX) throw Error(HSMD_ERR_ARG, "too many neighbour slots per site");
  general = disorder || j2ab_on;
  shellJ[0] = c.J1; shellJ[1] = c.J2; shellJ[2] = c.JX;
  shellLam[0] = lam(1); shellLam[1] = lam(2); shellLam[2] = lam(3);
  auto slot_col = [&](int s) {
    const int sh = slot_shell[s];
    return (sh == 1 ? 0 : sh == 2 ? zn1 : zn1 + zn2) + slot_k[s];
  };
  // colouring: sites sharing any active coupling get different colours.
  // First-fit in site order with a per-colour cap so classes stay balanced.
  {
    std::vector<std::vector<int32_t>> adj(N);
    for (int i = 0; i < N; i++)
      for (int s = 0; s < Z; s++) {
        const int nb = tab[(size_t)i * Zt + slot_col(s)];
        if (nb >= 0 && nb != i) { adj[i].push_back(nb); adj[nb].push_back(i); }
      }
    auto greedy = [&](int cap, std::vector<int32_t> &col) {
      col.assign(N, -1);
      std::vector<int> cnt;
      std::vector<char> used;
      for (int i = 0; i < N; i++) {
        used.assign(cnt.size() + 1, 0);
        for (int nb : adj[i])
          if (col[nb] >= 0 && col[nb] < (int)used.size()) used[col[nb]] = 1;
        int cc = 0;
        while (cc < (int)cnt.size() && (used[cc] || cnt[cc] >= cap)) cc++;
        if (cc == (int)cnt.size()) cnt.push_back(0);
        col[i] = cc;
        cnt[cc]++;
      }
      return (int)cnt.size();
    };
    std::vector<int32_t> col;
    int nc;
    if (col_cap > 0) {
      nc = greedy(col_cap, col);
    } else {
      nc = greedy(N, col);
      int cap = (N + nc - 1) / nc;
      std::vector<int32_t> col2;
      int nc2 = greedy(cap, col2);
      if (nc2 <= nc + 2) { col = col2; nc = nc2; }
      // classes larger than one workgroup's threads (1024) rule out the
      // MC kernel that keeps each site's labels in registers; for lattices
      // whose FP32 replica fits LDS (N <= 13000), where a cap of 1024 sites
      // per class stays within the colour limit, take it (irregular graphs:
      // more classes, each of one pass of the threads)
      if ((N + nc - 1) / nc > 1024 && N <= 13000) {
        std::vector<int32_t> col3;
        const int nc3 = greedy(1024, col3);
        if (nc3 <= CMAX - 2) { col = col3; nc = nc3; }
      }
    }
    if (nc > CMAX) throw Error(HSMD_ERR_ARG, "graph needs too many colours");
    colour = col;
    n_colours = nc;
  }
  // relabel: sort by (colour, coupled-neighbour count descending, original
  // index).  Sites of one colour are updated independently, so the order
  // inside a class changes no result; on irregular graphs it gives each
  // 64-site wave of a class neighbours counts close to one another.
  std::vector<int> degree(N, 0);
  for (int i = 0; i < N; i++)
    for (int s = 0; s < Z; s++) degree[i] += tab[(size_t)i * Zt + slot_col(s)] >= 0 ? 1 : 0;
  perm.resize(N);
  for (int i = 0; i < N; i++) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
    return colour[a] != colour[b] ? colour[a] < colour[b] : degree[a] > degree[b];
  });
  inv.assign(N, 0);
  for (int p = 0; p < N; p++) inv[perm[p]] = p;
  col_off.assign(n_colours + 1, 0);
  for (int i = 0; i < N; i++) col_off[colour[i] + 1]++;
  for (int cc = 0; cc < n_colours; cc++) col_off[cc + 1] += col_off[cc];
  g_perm = perm;
  // GPU neighbour table, slot-major, new labels, sentinel N
  g_nb.assign((size_t)Z * N, N);
  if (general) g_J.assign((size_t)Z * N, 0.0); else g_J.clear();
  for (int p = 0; p < N; p++) {
    const int i = perm[p];
    for (int s = 0; s < Z; s++) {
      const int nb = tab[(size_t)i * Zt + slot_col(s)];
      if (nb < 0) continue;
      g_nb[(size_t)s * N + p] = inv[nb];
      if (general) g_J[(size_t)s * N + p] = slotJ(slot_shell[s], slot_k[s], i, nb);
    }
  }
  // bonds with classes
  g_bi.resize(bi.size()); g_bj.resize(bi.size()); g_bc.resize(bi.size());
  bc_Jxy.clear(); bc_Jz.clear();
  std::map<std::pair<double, double>, int> cls;
  for (size_t b = 0; b < bi.size(); b++) {
    const double J = bondJ(bshell[b], bi[b], bj[b]);
    const double jxy = J, jz = J * lam(bshell[b]);
    auto key = std::make_pair(jxy, jz);
    auto it = cls.find(key);
    int id;
    if (it == cls.end()) {
      id = (int)bc_Jxy.size();
      if (id >= BCMAX) throw Error(HSMD_ERR_ARG, "too many distinct bond couplings");
      cls[key] = id;
      bc_Jxy.push_back(jxy); bc_Jz.push_back(jz);
    } else {
      id = it->second;
    }
    g_bi[b] = inv[bi[b]]; g_bj[b] = inv[bj[b]]; g_bc[b] = (uint8_t)id;
  }
}

bool Lattice::field_energy_ok() const {
  if (general) return false;
  // a fixed pseudo-random unit-vector configuration
  std::vector<double> S((size_t)3 * N);
  uint32_t st = 0x9e3779b9u;
  auto rnd = [&]() { st ^= st << 13; st ^= st >> 17; st ^= st << 5; return (st >> 8) * (1.0 / 16777216.0) - 0.5; };
  for (int i = 0; i < N; i++) {
    double v[3] = {rnd(), rnd(), rnd()};
    const double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + 1e-12;
    for (int d = 0; d < 3; d++) S[3 * i + d] = v[d] / n;
  }
  double eb = 0.0, ef = 0.0, sz = 0.0;
  for (size_t b = 0; b < bi.size(); b++) {
    const double J = bondJ(bshell[b], bi[b], bj[b]), Jz = J * lam(bshell[b]);
    const int i = bi[b], j = bj[b];
    eb += J * (S[3 * i] * S[3 * j] + S[3 * i + 1] * S[3 * j + 1]) + Jz * S[3 * i + 2] * S[3 * j + 2];
  }
  for (int p = 0; p < N; p++) {
    const int i = perm[p];
    double h[3] = {0, 0, 0};
    for (int s = 0; s < Z; s++) {
      const int q = g_nb[(size_t)s * N + p];
      if (q >= N) continue;
      const int j = perm[q];
      const int sh = slot_shell[s];
      h[0] += shellJ[sh - 1] * S[3 * j];
      h[1] += shellJ[sh - 1] * S[3 * j + 1];
      h[2] += shellJ[sh - 1] * shellLam[sh - 1] * S[3 * j + 2];
    }
    ef += 0.5 * (S[3 * i] * h[0] + S[3 * i + 1] * h[1] + S[3 * i + 2] * h[2]);
    sz += S[3 * i + 2];
  }
  (void)sz;
  return fabs(eb - ef) <= 1e-9 * (1.0 + fabs(eb));
}

}  // namespace hsmd
