// Text outputs of the MC static structure factor and of the SMD dynamic
// structure factor, written by the drop-in programs after the GPU work:
//   ordered spectrum arrays   get_OrderedSpecArray2D CF/subrouts.cpp:8868-8930
//   zone maps                 record_SpecArray2D     CF/subrouts.cpp:8979-9130
//   wavevector path           get_qPathWVectors      CF/subrouts.cpp:3688-3843
//   spectrum along the path   get_Spec_qPath         CF/subrouts.cpp:9165-9280
//   MC files                  CF/mc_hm_code.cpp:5752-5930
//   SMD files                 CF/smd_code.cpp:3628-4130
// Host-only I/O; the spectra come from the GPU engines.
#include <math.h>
#include <stdio.h>
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "hsmd_internal.h"

namespace hsmd {

namespace {

const double kPi = acos(-1.0);
const double kPi2 = 2.0 * kPi;
const double kS60 = sin(kPi / 3.0);
const double kQval = 4.0 * kPi / 3.0;   // CF/subrouts.cpp:594
// reciprocal vectors of the triangular Bravais lattice (CF/subrouts.cpp:571-573)
const double kB1[2] = {kPi2, -kPi / kS60};
const double kB2[2] = {0.0, kPi2 / kS60};

void make_dir(const std::string &p) { mkdir(p.c_str(), 0755); }

// std::scientific, showpos, precision dg, width wd (fmtDbleSci, CF/aux_code.cpp:545)
std::string fmt_sci(double v, int dg, int wd) {
  std::ostringstream o;
  o << std::scientific << std::showpos << std::setprecision(dg) << std::setw(wd) << v;
  return o.str();
}

std::string fixed3(double v) {
  std::ostringstream o;
  o << std::fixed << std::setprecision(3) << v;
  return o.str();
}

// hexagonal first-zone test (hexFilter, CF/aux_code.cpp:1189-1217)
bool hex_filter(double qx, double qy, double Q0, double dQ) {
  const double sq3 = sqrt(3.0), s60 = 0.5 * sq3;
  const double Qs = Q0 + dQ / 3.0;
  if (fabs(qy) - Qs * s60 < 0.0) {
    const double w1 = sq3 * (qx - Qs) - qy, w2 = sq3 * (qx - Qs) + qy;
    const double w3 = sq3 * (qx + Qs) - qy, w4 = sq3 * (qx + Qs) + qy;
    if ((w1 < 0.0 && w2 < 0.0) && (w3 > 0.0 && w4 > 0.0)) return true;
  }
  return false;
}

struct Grid {
  int G;
  bool c3;        // hexagonal zone (triangular, hexagonal, kagome) or square-like
  double dq;
};

// xy = (x + y)/2, yz = (y + z)/2, zz, tt = (x + y + z)/3 on the centred grid;
// spec is Nq x 3 in FFT order.
void ordered(const Grid &g, const double *spec, std::vector<double> (&o)[4]) {
  const int G = g.G, l2 = G / 2;
  std::vector<int> mv(G);
  for (int i = 0; i < G; i++) { const int n = 2 * (i / l2); mv[i] = i - l2 * (n - 1); }
  for (int f = 0; f < 4; f++) o[f].assign((size_t)G * G, 0.0);
  for (int i = 0; i < G; i++)
    for (int j = 0; j < G; j++) {
      const size_t k = (size_t)mv[i] * G + mv[j];
      const double x = spec[3 * k], y = spec[3 * k + 1], z = spec[3 * k + 2];
      const size_t t = (size_t)i * G + j;
      o[0][t] = 0.5 * (x + y);
      o[1][t] = 0.5 * (y + z);
      o[2][t] = z;
      o[3][t] = (1.0 / 3.0) * (x + y + z);
    }
}

// grid (i, j) pairs along the KGMYG (C3) or YGMXG path
std::vector<std::pair<int, int>> path_indices(const Grid &g) {
  const int G = g.G, l2 = G / 2;
  std::vector<std::pair<int, int>> p;
  if (g.c3) {
    for (int j = l2 / 3; j <= l2; j++) p.push_back({G - j, j});        // K -> G
    for (int j = l2 - 1; j >= 0; j--) p.push_back({l2, j});             // G -> M
    for (int j = 1; j <= l2 / 2; j++) p.push_back({l2 - j, j});         // M -> Y
    for (int i = l2 / 2 + 1; i <= l2; i++) p.push_back({i, i});         // Y -> G
  } else {
    for (int j = 0; j <= l2; j++) p.push_back({l2, j});                 // Y -> G
    for (int i = l2 - 1; i >= 0; i--) p.push_back({i, i});              // G -> M
    for (int j = 1; j <= l2; j++) p.push_back({0, j});                  // M -> X
    for (int i = 1; i <= l2; i++) p.push_back({i, l2});                 // X -> G
  }
  return p;
}

// cumulative length along the path wavevectors
std::vector<double> path_positions(const Grid &g) {
  const int G = g.G, l2 = G / 2;
  std::vector<std::pair<double, double>> w;
  if (g.c3) {
    const double b1[2] = {kB1[0] / G, kB1[1] / G}, b2[2] = {kB2[0] / G, kB2[1] / G};
    for (const auto &ij : path_indices(g)) {
      const int m1 = ij.first - l2, m2 = ij.second - l2;
      w.push_back({m1 * b1[0] + m2 * b2[0], m1 * b1[1] + m2 * b2[1]});
    }
  } else {
    const double dq = g.dq;
    for (int i = 0; i < l2; i++) w.push_back({0.0, i * dq - kPi});      // Y -> G
    std::vector<std::pair<double, double>> v;
    for (int j = l2; j >= 0; j--) v.push_back({j * dq - kPi, 0.0});      // G -> X
    for (int i = l2 - 1; i >= 0; i--) v.push_back({-kPi, i * dq - kPi}); // X -> M
    for (int i = 1; i <= l2; i++) v.push_back({i * dq - kPi, i * dq - kPi});  // M -> G
    std::reverse(v.begin(), v.end());
    w.insert(w.end(), v.begin(), v.end());
  }
  std::vector<double> pos(w.size());
  double acc = 0.0;
  std::pair<double, double> ref = w.empty() ? std::make_pair(0.0, 0.0) : w[0];
  for (size_t k = 0; k < w.size(); k++) {
    const double dx = w[k].first - ref.first, dy = w[k].second - ref.second;
    acc += sqrt(dx * dx + dy * dy);
    pos[k] = acc;
    ref = w[k];
  }
  return pos;
}

// zone map "qx qy value" rows (record_SpecArray2D)
void record_map(const std::string &file, const Grid &g, const std::vector<double> &a) {
  std::ofstream o(file);
  if (!o) throw Error(HSMD_ERR_IO, "cannot write " + file);
  const int sz = g.G, l2 = sz / 2;
  const std::string X3 = "   ";
  if (g.c3) {
    const double b1[2] = {kB1[0] / sz, kB1[1] / sz}, b2[2] = {kB2[0] / sz, kB2[1] / sz};
    const double q1[2] = {l2 * b1[0], l2 * b1[1]}, q2[2] = {l2 * b2[0], l2 * b2[1]};
    for (int n1 = 0; n1 < 3; n1++) {
      for (int i = 0; i < sz; i++) {
        for (int n2 = 0; n2 < 3; n2++) {
          const int m1s = 2 * (n1 - 1), m2s = 2 * (n2 - 1);
          const double qx0 = m1s * q1[0] + m2s * q2[0], qy0 = m1s * q1[1] + m2s * q2[1];
          for (int j = 0; j < sz; j++) {
            const double pw = a[(size_t)i * sz + j];
            const int m1 = i - l2, m2 = j - l2;
            const double qx = qx0 + m1 * b1[0] + m2 * b2[0], qy = qy0 + m1 * b1[1] + m2 * b2[1];
            if (hex_filter(qx, qy, kQval, g.dq)) o << qx << X3 << qy << X3 << pw << "\n";
            else o << qx << X3 << qy << X3 << "NaN" << "\n";
          }
        }
        o << "\n";
      }
    }
  } else {
    for (int i = 0; i <= sz; i++) {
      for (int j = 0; j <= sz; j++) {
        const int ii = i == sz ? 0 : i, jj = j == sz ? 0 : j;
        const double pw = a[(size_t)ii * sz + jj];
        o << g.dq * (i - l2) << X3 << g.dq * (j - l2) << X3 << pw << "\n";
      }
      o << "\n";
    }
  }
}

Grid grid_of(const Lattice &L) {
  Grid g;
  g.G = L.Gsz;
  g.c3 = L.c3sym;
  g.dq = L.sq_dq();
  if (g.G < 2 || g.G % 2) throw Error(HSMD_ERR_ARG, "spectrum outputs need an even grid side");
  return g;
}

}  // namespace

// MC: DFT_{xy,yz,zz,tt}Spec zone maps of the averaged static S(q) and, for
// generated lattices, their values along the path (*_wZero_DFT_*Spec).
// sq: Nq x 3 complex averages (the reference's Sqx0/Sqy0/Sqz0 after the run).
void write_mc_spectra(const std::string &dir, const Lattice &L, const double *sq,
                      const std::string &tagDat) {
  const Grid g = grid_of(L);
  const int Nq = L.Nq;
  std::vector<double> spec((size_t)3 * Nq);
  for (int k = 0; k < Nq; k++)
    for (int d = 0; d < 3; d++) {
      const double re = sq[((size_t)k * 3 + d) * 2], im = sq[((size_t)k * 3 + d) * 2 + 1];
      spec[(size_t)3 * k + d] = std::max(1.0e-16, re * re + im * im);
    }
  std::vector<double> o[4];
  ordered(g, spec.data(), o);
  const std::string base = dir + "/outData/";
  make_dir(base);
  make_dir(base + "MC_Results");
  const char *nm[4] = {"xy", "yz", "zz", "tt"};
  for (int f = 0; f < 4; f++)
    record_map(base + "MC_Results/DFT_" + nm[f] + "Spec" + tagDat, g, o[f]);
  if (L.tabulated) return;
  const auto ij = path_indices(g);
  const auto lp = path_positions(g);
  const std::string pre = g.c3 ? "KGMYG" : "YGMXG";
  for (int f = 0; f < 4; f++) {
    std::ofstream out(base + "MC_Results/" + pre + "_wZero_DFT_" + nm[f] + "Spec" + tagDat);
    for (size_t k = 0; k < ij.size(); k++)
      out << fmt_sci(lp[k], 8, 15) << fmt_sci(o[f][(size_t)ij[k].first * g.G + ij[k].second], 12, 22)
          << "\n";
  }
}

// SMD: quantum-corrected spectra (beta w / (1 - e^{-beta w}) x tmax / 2 pi),
// w-slices every ceil(0.1 npw) bins, the path files over the npw = ntm / 2
// positive frequencies (and the zz cut), and SpinSpin_Temporal_Corr.
// dssf: Nq x 3 x ntm sample averages (the MC_Bin_SpecData layout), qss: ntm.
// The reference's y component is lost to an uninitialised variable
// (CF/smd_code.cpp:3662-3667); here y is corrected like x and z.
void write_smd_spectra(const std::string &dir, const Lattice &L, const double *dssf,
                       const double *qss, int ntm, double dtm, double Temp,
                       const std::string &tagDat) {
  const Grid g = grid_of(L);
  const int Nq = L.Nq, npw = ntm / 2;
  const double tmax = ntm * dtm, dwf = kPi2 / tmax, beta = 1.0 / Temp, itFac = tmax / kPi2;
  std::vector<double> c((size_t)Nq * 3 * ntm);
  for (int k = 0; k < Nq; k++)
    for (int n = 0; n < ntm; n++) {
      const double fc = beta * n * dwf;
      const double qt = n > 0 ? fc / (1.0 - exp(-fc)) : 1.0;
      for (int d = 0; d < 3; d++)
        c[((size_t)k * 3 + d) * ntm + n] = itFac * qt * dssf[((size_t)k * 3 + d) * ntm + n];
    }
  const std::string base = dir + "/outData/";
  make_dir(base);
  make_dir(base + "MC_Results");
  make_dir(base + "RK_Results");
  const char *nm[4] = {"xy", "yz", "zz", "tt"};
  const auto ij = path_indices(g);
  const size_t np = ij.size();
  std::vector<double> spec((size_t)Nq * 3), path[4];
  for (int f = 0; f < 4; f++) path[f].assign(np * npw, 0.0);
  std::vector<double> o[4];
  const int nmax = npw;
  const int nwSkip = std::max(1, (int)ceil(0.1 * nmax));
  for (int n = 0; n < npw; n++) {
    for (int k = 0; k < Nq; k++)
      for (int d = 0; d < 3; d++) spec[(size_t)3 * k + d] = c[((size_t)k * 3 + d) * ntm + n];
    ordered(g, spec.data(), o);
    for (int f = 0; f < 4; f++)
      for (size_t p = 0; p < np; p++)
        path[f][p * npw + n] = o[f][(size_t)ij[p].first * g.G + ij[p].second];
    if (n % nwSkip == 0) {
      const std::string w = fixed3(n * dwf);
      for (int f = 0; f < 4; f++)
        record_map(base + "RK_Results/DFT_" + nm[f] + "Spec_wfreq(" + w + ")" + tagDat, g, o[f]);
    }
  }
  {
    std::ofstream out(base + "MC_Results/SpinSpin_Temporal_Corr" + tagDat);
    for (int n = 0; n < ntm; n++) out << fmt_sci(n * dtm, 8, 15) << "    " << fmt_sci(qss[n], 12, 22) << "\n";
  }
  const auto lp = path_positions(g);
  const bool sq_like = !g.c3;
  const std::string pre = sq_like ? "YGMXG" : "KGMYG";
  const double cut1 = sq_like ? 0.5 * kPi : 0.0, cut2 = sq_like ? kPi * (1.0 + 0.5 * sqrt(2.0)) : kPi;
  std::ofstream out[4], cut;
  for (int f = 0; f < 4; f++) out[f].open(base + "RK_Results/" + pre + "_wDepn_DFT_" + nm[f] + "Spec" + tagDat);
  cut.open(base + "RK_Results/" + std::string(sq_like ? "YGMCut" : "KGMCut") + "_wDepn_DFT_zzSpec" + tagDat);
  for (size_t k = 0; k < np; k++) {
    const double l = lp[k];
    for (int n = 0; n < nmax; n++) {
      const double w = n * dwf;
      for (int f = 0; f < 4; f++)
        out[f] << fmt_sci(l, 8, 15) << "    " << fmt_sci(w, 8, 15) << "    "
               << fmt_sci(path[f][k * npw + n], 12, 22) << "\n";
      if (l >= cut1 && l <= cut2)
        cut << fmt_sci(l, 8, 15) << "    " << fmt_sci(w, 8, 15) << "    "
            << fmt_sci(path[2][k * npw + n], 12, 22) << "\n";
    }
    for (int f = 0; f < 4; f++) out[f] << "\n";
    cut << "\n";
  }
}

// Quasicrystal histograms of one temperature (CF/mc_hm_code.cpp:5961-6130):
// counts h[HSMD_QC_HIST] (Neel [41], stripes [41][4], stripe vector
// [41 imag][41 real]) added to the counts already in the binary files
// MC_Bin_NeelHist / StrpHist / SMagVecHist<tag>.bin (a file of another size
// is an error, as in the reference), written back, and the normalised
// histograms as outData/MC_Results/NeelHist / StrpHist / SMagVecHist<tag>.dat.
void write_qc_histograms(const std::string &dir, const int32_t *h, const std::string &tag) {
  constexpr int H1 = 41;
  const double box = 2.0 / (H1 - 1);
  std::vector<int32_t> neel(h, h + H1), strp(h + H1, h + 5 * H1), vec(h + 5 * H1, h + 5 * H1 + H1 * H1);
  auto merge = [&](const std::string &name, std::vector<int32_t> &v) {
    const std::string path = dir + "/outBin/" + name + tag + ".bin";
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (in) {
      const long long sz = (long long)in.tellg();
      if (sz != (long long)(v.size() * sizeof(int32_t)))
        throw Error(HSMD_ERR_IO, "histogram file of the wrong size: " + path);
      in.seekg(0);
      std::vector<int32_t> old(v.size());
      in.read(reinterpret_cast<char *>(old.data()), sz);
      for (size_t i = 0; i < v.size(); i++) v[i] += old[i];
    }
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) throw Error(HSMD_ERR_IO, "cannot write " + path);
    out.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(int32_t)));
  };
  merge("MC_Bin_NeelHist", neel);
  merge("MC_Bin_StrpHist", strp);
  merge("MC_Bin_SMagVecHist", vec);
  auto fix = [](double x) {
    char b[64];
    snprintf(b, sizeof b, "%+8.4f", x);
    return std::string(b);
  };
  long long n0 = 0, n4[4] = {0, 0, 0, 0}, nv = 0;
  for (int i = 0; i < H1; i++) {
    n0 += neel[i];
    for (int k = 0; k < 4; k++) n4[k] += strp[4 * i + k];
  }
  for (int v : vec) nv += v;
  const std::string base = dir + "/outData/MC_Results/";
  make_dir(dir + "/outData/");
  make_dir(base);
  std::ofstream o1(base + "NeelHist" + tag + ".dat"), o2(base + "StrpHist" + tag + ".dat");
  if (!o1 || !o2) throw Error(HSMD_ERR_IO, "cannot write histogram files in " + base);
  for (int i = 0; i < H1; i++) {
    const double x = i * box - 1.0;
    o1 << fix(x) << "  " << fmt_sci(neel[i] / (double)n0, 6, 14) << "  \n";
    o2 << fix(x) << "  ";
    for (int k = 0; k < 4; k++) o2 << fmt_sci(strp[4 * i + k] / (double)n4[k], 6, 14) << "  ";
    o2 << "\n";
  }
  std::ofstream o3(base + "SMagVecHist" + tag + ".dat");
  if (!o3) throw Error(HSMD_ERR_IO, "cannot write histogram files in " + base);
  for (int i = 0; i < H1; i++) {
    for (int j = 0; j < H1; j++)
      o3 << fix(i * box - 1.0) << "  " << fix(j * box - 1.0) << "  "
         << fmt_sci(vec[(size_t)j * H1 + i] / (double)nv, 6, 14) << "\n";
    o3 << "\n";
  }
}

}  // namespace hsmd
