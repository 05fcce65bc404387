// Drop-in host drivers: the MC and SMD runs of the reference, reading the
// same Settings_*.txt files (written by Set_Params.py) from the working
// directory and writing the same outBin/ binary files, with the hot loops
// running through the hsmd C ABI.
//
// Reference drivers: main of CF/mc_hm_code.cpp:139 (xMC_run) and of
// CF/smd_code.cpp:141 (xSMD_run).  Settings parsing and guard rails follow
// convert_inputs2data (CF/subrouts.cpp:955-2186); file names follow
// CF/mc_hm_code.cpp:2209-2214, 3944-3972 and CF/smd_code.cpp:905-930.
#include <math.h>
#include <stdio.h>
#include <sys/stat.h>

#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hsmd.h"
#include "hsmd_common.h"
#include "hsmd_internal.h"

namespace hsmd {

// postproc.cpp
void write_qc_histograms(const std::string &dir, const int32_t *h, const std::string &tag);
void write_mc_spectra(const std::string &dir, const Lattice &L, const double *sq,
                      const std::string &tagDat);
void write_smd_spectra(const std::string &dir, const Lattice &L, const double *dssf,
                       const double *qss, int ntm, double dtm, double Temp,
                       const std::string &tagDat);

// Temporal windows of the DSSF time transform (CF/smd_code.cpp:2808-2895), on
// n points: gauss exp(-80 x^2) and the sinc windows on x in [-1, 1]; Hann^p
// and the cosine-sum windows (Nuttall, Blackman-Harris, flat top) on
// x = i / (n - 1).  Returns false for a name it does not know (the file-read
// "wread" is the caller's).
static bool temporal_window(const std::string &f, int n, double *w) {
  const double tp = 2.0 * acos(-1.0);
  if (f == "nowin") {
    for (int i = 0; i < n; i++) w[i] = 1.0;
  } else if (f == "gauss") {
    const double dx = 2.0 / (n - 1.0);
    for (int i = 0; i < n; i++) { const double x = i * dx - 1.0; w[i] = exp(-80.0 * pow(x, 2)); }
  } else if (f == "lancz" || f == "sincw") {
    const double fc = f == "lancz" ? 0.5 : 1.0, dx = 2.0 / (n - 1.0);
    for (int i = 0; i < n; i++) {
      const double x = (i * dx - 1.0) * fc;
      const double a = tp * x;
      w[i] = (f == "lancz" && fabs(x) > 0.5) ? 0.0 : (a == 0.0 ? 1.0 : sin(a) / a);
    }
  } else if (f == "hann1" || f == "hann2" || f == "hann3") {
    const int pw = f[4] - '0';
    const double dx = 1.0 / (n - 1.0);
    for (int i = 0; i < n; i++) w[i] = pow(0.5 * (1.0 - cos(tp * i * dx)), pw);
  } else if (f == "black" || f == "ntall" || f == "flatw") {
    const double nut[5] = {0.3635819, 0.4891775, 0.1365995, 0.0106411, 0};
    const double har[5] = {0.35875, 0.48829, 0.14128, 0.01168, 0};
    const double flt[5] = {0.215578950, 0.416631580, 0.277263158, 0.083578947, 0.006947368};
    const double *a = f == "ntall" ? nut : f == "black" ? har : flt;
    const double dx = 1.0 / (n - 1.0);
    for (int i = 0; i < n; i++) {
      const double x = i * dx;
      w[i] = a[0] + (-a[1] * cos(1.0 * tp * x) + a[2] * cos(2.0 * tp * x) - a[3] * cos(3.0 * tp * x) +
                     a[4] * cos(4.0 * tp * x));
    }
  } else {
    return false;
  }
  return true;
}

namespace {

// Settings_CodeExec.txt order (Set_Params.py code_exec_labels):
// ptON ptAdapt ptTrack ANNL_ON FFTW_ON RKCheck PBC_OFF IsiModel qcrystal
// with_recField with_DFTcodes force_Stripes refState_Zero qct_NeelPhase
// qct_RemovePBC rec_orderMaps
enum { PT_ON, PT_ADAPT, PT_TRACK, ANNL_ON, FFTW_ON, RK_CHECK, PBC_OFF, ISI_MODEL, QCRYSTAL,
       WITH_RECFIELD, WITH_DFT, FORCE_STRIPES, REFSTATE_ZERO, QCT_NEELPHASE };

constexpr int kBinSz = 1000;        // CF/subrouts.cpp:309
constexpr double kTempMax = 2.0;    // CF/subrouts.cpp:373
constexpr double kDwfMin = 0.10;    // CF/subrouts.cpp:391

struct Settings {
  std::string dir;
  // Settings_MODEL.txt
  std::string size_tok, geom;
  int L = 0;
  double J1 = 0, J2 = 0, JX = 0, J2d = 0, l1 = 1, l2 = 1, h = 0, T1 = 0, T2 = 0;
  // Settings_MC_SIM.txt
  int NTerm = 0, NMeas = 0, crSkip = 1, vrSkip = 0, vision = 0;
  std::string IState, TpGrid, outTag;
  // Settings_SMD_SIM.txt
  int ntm = 0, tpLstSz = 0, tpIndex = 0, rec1 = 0, rec2 = 0;
  double dtm = 0;
  std::string tWForm;
  // Settings_Disorder.txt
  double imp = 0, J1ic = 0, J2ic = 0, JXic = 0, J1ii = 0, J2ii = 0, JXii = 0;
  // Settings_CodeExec.txt
  int code[16] = {0};
  std::string label;  // outLabel, e.g. "_J2(0.600)"
};

[[noreturn]] void fail(int code, const std::string &m) { throw Error(code, m); }

std::vector<std::string> tokens(const std::string &path, size_t n) {
  std::ifstream f(path);
  if (!f.is_open()) fail(HSMD_ERR_IO, "Unable to open the input file: " + path);
  std::vector<std::string> v(n);
  for (size_t i = 0; i < n; i++)
    if (!(f >> v[i])) fail(HSMD_ERR_IO, "Input file has fewer than " + std::to_string(n) + " values: " + path);
  return v;
}

bool is_int(const std::string &s) {
  if (s.empty()) return false;
  size_t i = (s[0] == '-' || s[0] == '+') ? 1 : 0;
  if (i == s.size()) return false;
  for (; i < s.size(); i++) if (s[i] < '0' || s[i] > '9') return false;
  return true;
}

double to_f(const std::string &s, const char *what) {
  char *end = nullptr;
  const double v = strtod(s.c_str(), &end);
  if (end == s.c_str() || *end) fail(HSMD_ERR_GUARD, std::string("invalid number for ") + what + ": " + s);
  return v;
}

int to_i(const std::string &s, const char *what) {
  if (!is_int(s)) fail(HSMD_ERR_GUARD, std::string("invalid integer for ") + what + ": " + s);
  return atoi(s.c_str());
}

std::string fixed(double v, int p) {
  std::ostringstream o;
  o << std::fixed << std::setprecision(p) << v;
  return o.str();
}

// Reads and checks the five settings files (convert_inputs2data).
Settings read_settings(const std::string &dir, bool strict) {
  Settings s;
  s.dir = dir;
  {
    std::ifstream f(dir + "/Settings_CodeExec.txt");
    if (!f.is_open()) fail(HSMD_ERR_IO, "Unable to open Settings_CodeExec.txt");
    std::string k;
    int g;
    for (int i = 0; i < 16; i++) {
      if (!(f >> k >> g)) fail(HSMD_ERR_IO, "invalid setting pair (" + std::to_string(i) + ")");
      if (g != 0 && g != 1) fail(HSMD_ERR_GUARD, "code-exec setting must be 0 or 1: " + k);
      s.code[i] = g;
    }
  }
  auto m = tokens(dir + "/Settings_MODEL.txt", 11);
  auto c = tokens(dir + "/Settings_MC_SIM.txt", 8);
  auto d = tokens(dir + "/Settings_SMD_SIM.txt", 7);
  auto q = tokens(dir + "/Settings_Disorder.txt", 7);
  s.size_tok = m[0];
  s.J1 = to_f(m[1], "J1"); s.J2 = to_f(m[2], "J2"); s.JX = to_f(m[3], "JX");
  s.J2d = to_f(m[4], "J2_delta"); s.l1 = to_f(m[5], "lambda1"); s.l2 = to_f(m[6], "lambda2");
  s.h = to_f(m[7], "extH"); s.T1 = to_f(m[8], "Temp1"); s.T2 = to_f(m[9], "Temp2");
  s.geom = m[10];
  const bool qc = s.code[QCRYSTAL];
  if (qc) {
    if (s.geom != "Aamman") fail(HSMD_ERR_GUARD, "Unknown quasi-geometry : " + s.geom);
  } else if (s.geom != "square" && s.geom != "triang" && s.geom != "hexagn" &&
             s.geom != "kagome" && s.geom != "lieb") {
    fail(HSMD_ERR_GUARD, "Unknown geometry : " + s.geom);
  }
  s.L = to_i(s.size_tok, "lattice size");
  const bool c3 = !qc && s.geom != "square" && s.geom != "lieb";
  if (!qc && s.L % 2 != 0) fail(HSMD_ERR_GUARD, "Lattice length must multiple of 2!");
  if (!qc && c3 && s.L % 3 != 0 && strict) fail(HSMD_ERR_GUARD, "Lattice length must multiple of 3!");
  if (fabs(s.J2d) > 1e-12 && s.geom != "lieb")
    fail(HSMD_ERR_GUARD, "AM d-wave J2-coupling requires the Lieb lattice (change inputs).");
  if (s.code[ISI_MODEL] && fabs(s.h) > 1e-12 && strict)
    fail(HSMD_ERR_GUARD, "Ext. magnetic field must point in the x-direction (Ising model)!");
  if (s.T1 <= 0) fail(HSMD_ERR_GUARD, "Invalid temperature (1);");
  if (s.T2 < s.T1) fail(HSMD_ERR_GUARD, "Invalid temperature (2);");
  s.NTerm = to_i(c[0], "NTerm");
  s.NMeas = to_i(c[1], "NMeas");
  if (s.NMeas % kBinSz != 0) fail(HSMD_ERR_GUARD, "Number of measurements must be a multiple of the binning size : BinSz = 1000");
  if (s.NMeas / kBinSz == 1) fail(HSMD_ERR_GUARD, "Number of bins must be greater than 1!");
  s.IState = c[2]; s.TpGrid = c[3]; s.outTag = c[4];
  const char *states[] = {"loww_T", "high_T", "zpolar", "2up1dw", "hstrip", "vstrip", "inputf"};
  bool okst = false;
  for (const char *x : states) okst |= (s.IState == x);
  if (!okst) fail(HSMD_ERR_GUARD, "Initial state string is invalid (check inputs);");
  const std::string g2 = qc ? "square" : s.geom;
  if (g2 == "square" && s.IState == "2up1dw") fail(HSMD_ERR_GUARD, "Initial state string is not compatible with the chosen system geometry;");
  if (g2 != "square" && g2 != "triang" && s.IState != "high_T" && s.IState != "zpolar" && s.IState != "inputf")
    fail(HSMD_ERR_GUARD, "Initial state string is not compatible with the chosen system geometry;");
  if (s.code[ISI_MODEL] && s.IState != "high_T" && s.IState != "zpolar" && s.IState != "inputf")
    fail(HSMD_ERR_GUARD, "Initial state string is not compatible with the chosen spin-model type;");
  if (s.TpGrid != "linear" && s.TpGrid != "nonlin" && s.TpGrid != "custom")
    fail(HSMD_ERR_GUARD, "Invalid T-grid option, (please check inputs);");
  if (s.outTag != "Htag" && s.outTag != "J1tag" && s.outTag != "J2tag")
    fail(HSMD_ERR_GUARD, "Invalid output label option, (please check and fix inputs);");
  s.crSkip = to_i(c[5], "crSkip");
  if (s.crSkip <= 0 || kBinSz % s.crSkip || s.crSkip > kBinSz)
    fail(HSMD_ERR_GUARD, "Spin configuration recording skip number has to be a divisor of BinSz;");
  s.vrSkip = to_i(c[6], "vrSkip");
  s.vision = to_i(c[7], "MClive");
  s.ntm = to_i(d[0], "ntm");
  if (s.ntm < 10) fail(HSMD_ERR_GUARD, "Number of time steps for the evolution RK-procedure is too small");
  s.tpLstSz = to_i(d[1], "tpLstSz");
  if (s.tpLstSz < 1) fail(HSMD_ERR_GUARD, "Input temperature list size must be positive integer...");
  s.tpIndex = to_i(d[2], "tpIndex");
  if (s.tpIndex < 0 || s.tpIndex > s.tpLstSz - 1) fail(HSMD_ERR_GUARD, "Input temp. list position has to be smaller than tpLstSz");
  s.dtm = to_f(d[3], "dtm");
  if (s.dtm > 0.05) fail(HSMD_ERR_GUARD, "Temporal discretization is too large to keep energy conserved");
  s.tWForm = d[4];
  const char *wins[] = {"wread", "gauss", "ntall", "black", "flatw", "hann1", "sincw", "hann2",
                        "lancz", "hann3", "nowin"};
  bool okw = false;
  for (const char *x : wins) okw |= (s.tWForm == x);
  if (!okw) fail(HSMD_ERR_GUARD, "Invalid temporal window (please check inputs);");
  s.rec1 = to_i(d[5], "RecVd1"); s.rec2 = to_i(d[6], "RecVd2");
  s.imp = to_f(q[0], "impurity ratio");
  if (s.imp < 0.0 || s.imp >= 1.0) fail(HSMD_ERR_GUARD, "Impurity ratio must be a number in the range [0,1);");
  s.J1ic = to_f(q[1], "J1ic"); s.J2ic = to_f(q[2], "J2ic"); s.JXic = to_f(q[3], "JXic");
  s.J1ii = to_f(q[4], "J1ii"); s.J2ii = to_f(q[5], "J2ii"); s.JXii = to_f(q[6], "JXii");
  if (s.code[FFTW_ON]) fail(HSMD_ERR_GUARD, "FFTW_ON is not supported (the reference aborts on it too)");
  if (s.code[PT_ADAPT] || s.code[PT_TRACK] || s.code[ANNL_ON])
    fail(HSMD_ERR_GUARD, "ptAdapt / ptTrack / ANNL_ON are not built (experimental in the reference)");
  double lv;
  if (s.outTag == "Htag") { s.label = "_extH("; lv = s.h; }
  else if (s.outTag == "J2tag") { s.label = "_J2("; lv = s.J2; }
  else { s.label = "_J1("; lv = s.J1; }
  s.label += fixed(lv, 3) + ")";
  return s;
}

hsmd_couplings couplings_of(const Settings &s) {
  hsmd_couplings c;
  c.J1 = s.J1; c.J2 = s.J2; c.JX = s.JX; c.J2_delta = s.J2d;
  c.lambda1 = s.l1; c.lambda2 = s.l2; c.h = s.h; c.imp_ratio = s.imp;
  c.J1ic = s.J1ic; c.J2ic = s.J2ic; c.JXic = s.JXic;
  c.J1ii = s.J1ii; c.J2ii = s.J2ii; c.JXii = s.JXii;
  return c;
}

void check(hsmd_status st) {
  if (st != HSMD_OK) {
    char buf[1024];
    hsmd_last_error(buf, sizeof(buf));
    throw Error(st, buf);
  }
}

struct LatHolder {
  hsmd_lattice *p = nullptr;
  ~LatHolder() { if (p) hsmd_lattice_destroy(p); }
};

bool file_size(const std::string &p, long long &sz) {
  struct stat st;
  if (stat(p.c_str(), &st) != 0) return false;
  sz = (long long)st.st_size;
  return true;
}

template <typename T>
std::vector<T> read_bin(const std::string &p, size_t n) {
  std::ifstream f(p, std::ios::binary);
  if (!f.is_open()) fail(HSMD_ERR_IO, "File not found: " + p);
  std::vector<T> v(n);
  f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(T)));
  if (!f) fail(HSMD_ERR_IO, "Invalid data in: " + p);
  return v;
}

template <typename T>
void write_bin(const std::string &p, const T *v, size_t n, bool append) {
  std::ofstream f(p, std::ios::binary | (append ? std::ios::app : std::ios::trunc));
  if (!f.is_open()) fail(HSMD_ERR_IO, "cannot write " + p);
  f.write(reinterpret_cast<const char *>(v), (std::streamsize)(n * sizeof(T)));
}

// Lattice of the settings: generated, or read from the quasicrystal tables
// (sitesList / nborsList with the _AppxtSz(N) tag, CF/mc_hm_code.cpp:955-1020,
// 1273-1320).  The generated table is also written to outBin/nborsList.bin
// as make_nborsTable does (CF/subrouts.cpp:3294-3316).
void make_lattice(const Settings &s, bool strict, LatHolder &lh) {
  const hsmd_couplings c = couplings_of(s);
  if (s.code[QCRYSTAL]) {
    const int n = s.L;
    const std::string tag = "_AppxtSz(" + std::to_string(n) + ")";
    auto rv = read_bin<double>(s.dir + "/outBin/sitesList" + tag + ".bin", (size_t)n * 2);
    long long sz;
    const std::string nbf = s.dir + "/outBin/nborsList" + tag + ".bin";
    if (!file_size(nbf, sz) || sz != (long long)n * 19 * 4) fail(HSMD_ERR_IO, "Invalid data in: " + nbf);
    auto nb = read_bin<int32_t>(nbf, (size_t)n * 19);
    check(hsmd_lattice_from_tables(n, 8, 3, 8, nb.data(), rv.data(), &c, &lh.p));
  } else {
    check(hsmd_lattice_generate(s.geom.c_str(), s.L, s.code[PBC_OFF] ? 0 : 1, strict ? 1 : 0, &c, &lh.p));
    int32_t info[16];
    check(hsmd_lattice_info(lh.p, info));
    const int N = info[0], Zt = info[1] + info[2] + info[3];
    std::vector<int32_t> nb((size_t)N * Zt);
    check(hsmd_lattice_tables(lh.p, nb.data(), nullptr));
    const std::string nbf = s.dir + "/outBin/nborsList.bin";
    long long sz;
    if (file_size(nbf, sz) && sz == (long long)nb.size() * 4) {
      auto old = read_bin<int32_t>(nbf, nb.size());
      if (old != nb) fail(HSMD_ERR_IO, "outBin/nborsList.bin does not match the lattice settings");
    } else {
      write_bin(nbf, nb.data(), nb.size(), false);
    }
  }
}

// Impurity map of make_impurityField (CF/subrouts.cpp:3321-3352) over a
// Fisher-Yates order drawn from the counter-based stream.
std::vector<int32_t> impurity_map(int N, double ratio, uint64_t seed) {
  std::vector<int32_t> order(N), imp(N, 0);
  for (int i = 0; i < N; i++) order[i] = i;
  for (int i = N - 1; i > 0; --i) {
    const u32x4 r = draw4(seed, (uint32_t)i, 0u, 0u, STREAM_INIT + 16u);
    const int j = (int)(r.x % (uint32_t)(i + 1));
    std::swap(order[i], order[j]);
  }
  int i0 = 1;
  for (int i = 0; i < N; i++) {
    imp[order[i]] = i0;
    if ((double)(i + 2) / N > ratio) i0 = 0;
  }
  return imp;
}

// Initial spin field of one replica (set_initialSpinField CF/subrouts.cpp:4670-4910).
void initial_state(const Settings &s, int L, int N, bool c3, int rank, uint64_t seed, double *S) {
  auto u = [&](int site, int k) {
    const u32x4 r = draw4(seed, (uint32_t)site, (uint32_t)k, (uint32_t)rank, STREAM_INIT);
    return std::make_pair(u01_f64(r.x, r.y), u01_f64(r.z, r.w));
  };
  const double pi = acos(-1.0);
  for (int k = 0; k < 3 * N; k++) S[k] = 0.0;
  if (s.IState == "inputf") {
    // CF/mc_hm_code.cpp:1937-2025: tabulated graphs start from the reference
    // stripe maps stateStrp{1..4} (x component, rank % 4); generated
    // lattices from outBin/inputSConf.bin (N x 3 f64) for every rank
    std::string f;
    if (s.code[QCRYSTAL])
      f = s.dir + "/outBin/stateStrp" + std::to_string(1 + rank % 4) + "_AppxtSz(" +
          std::to_string(N) + ").bin";
    else
      f = s.dir + "/outBin/inputSConf.bin";
    long long sz = 0;
    if (!file_size(f, sz) || sz != (long long)N * 24) fail(HSMD_ERR_IO, "Error: invalid input spin file " + f);
    const std::vector<double> v = read_bin<double>(f, (size_t)3 * N);
    for (int k = 0; k < N; k++) {
      if (s.code[QCRYSTAL]) S[3 * k] = v[3 * k];
      else for (int d = 0; d < 3; d++) S[3 * k + d] = v[3 * k + d];
    }
    // the reference aborts on states whose average norm is not 1
    double nrm = 0.0;
    for (int k = 0; k < N; k++) nrm += sqrt(S[3 * k] * S[3 * k] + S[3 * k + 1] * S[3 * k + 1] + S[3 * k + 2] * S[3 * k + 2]);
    if (fabs(nrm / N - 1.0) > 1.0e-6) fail(HSMD_ERR_IO, "Error: input spin state is not normalised: " + f);
    return;
  }
  if (s.code[ISI_MODEL]) {
    for (int k = 0; k < N; k++) {
      if (s.IState == "zpolar") S[3 * k] = 1.0;
      else S[3 * k] = (2.0 * u(k, 0).first - 1.0 >= 0.0) ? 1.0 : -1.0;
    }
    return;
  }
  // pltSeq[0] (CF/subrouts.cpp:2102-2121): z with a field or an easy-axis
  // XXZ average, else x
  const bool xxz = (s.l1 != 1.0) || (s.l2 != 1.0);
  const int m = (fabs(s.h) > 1e-12 || (xxz && 0.5 * (s.l1 + s.l2) > 1.0)) ? 2 : 0;
  auto phase = [&](const std::string &kind, int k) {
    const int i = k / L, j = k % L;
    if (kind == "H") return cos(j * pi);
    if (kind == "V") return cos(i * pi);
    if (kind == "D") return cos(pi * (i + j));
    return cos(((i + j + i + 2) % 3) * pi);  // 2-up 1-down (make_2U1DPhaseList)
  };
  if (s.IState == "loww_T") {
    if (s.J2 < s.J1 / 2) {
      if (c3) {
        const int m1 = m, m2 = 1;
        for (int k = 0; k < N && k < L * L; k++) {
          const int i = k / L, j = k % L;
          const double th = (4.0 * pi / 3.0) * (i + j * cos(pi / 3.0));
          S[3 * k + m1] = cos(th);
          S[3 * k + m2] = sin(th);
        }
      } else {
        for (int k = 0; k < N && k < L * L; k++) S[3 * k + m] = phase("D", k);
      }
    } else {
      const int n0 = c3 ? 3 : 2;
      const char *kind = (rank % n0 == 0) ? "H" : (rank % n0 == 1) ? "V" : "D";
      for (int k = 0; k < N && k < L * L; k++) S[3 * k + m] = phase(kind, k);
    }
  } else if (s.IState == "2up1dw") {
    for (int k = 0; k < N && k < L * L; k++) S[3 * k + m] = phase("U", k);
  } else if (s.IState == "hstrip") {
    for (int k = 0; k < N && k < L * L; k++) S[3 * k + m] = phase("H", k);
  } else if (s.IState == "vstrip") {
    for (int k = 0; k < N && k < L * L; k++) S[3 * k + m] = phase("V", k);
  } else if (s.IState == "zpolar") {
    for (int k = 0; k < N; k++) S[3 * k + 2] = 1.0;
  }
  const bool random = (s.IState == "high_T");
  for (int k = 0; k < N; k++) {
    const auto r = u(k, 1);
    const double tht = 2.0 * pi * r.first, phi = acos(1.0 - 2.0 * r.second);
    const double rv[3] = {sin(phi) * cos(tht), sin(phi) * sin(tht), cos(phi)};
    if (random) {
      for (int d = 0; d < 3; d++) S[3 * k + d] = rv[d];
    } else {
      // 5 % random-vector noise on the ordered state, then renormalise
      double v[3], n2 = 0;
      for (int d = 0; d < 3; d++) { v[d] = S[3 * k + d] + 0.05 * rv[d]; n2 += v[d] * v[d]; }
      const double nr = sqrt(n2);
      for (int d = 0; d < 3; d++) S[3 * k + d] = v[d] / nr;
    }
  }
}

std::vector<double> temperature_grid(const Settings &s, int npt) {
  std::vector<double> T(npt);
  const std::string tfile = s.dir + "/outBin/Temp_List" + s.label + ".bin";
  if (npt == 1) { T[0] = s.T1; return T; }
  const double d0 = (s.T2 - s.T1) / (npt - 1);
  if (s.TpGrid == "custom") {
    long long sz;
    if (!file_size(tfile, sz) || sz < (long long)npt * 8) fail(HSMD_ERR_IO, "T-grid file missing or short: " + tfile);
    T = read_bin<double>(tfile, npt);
  } else if (s.TpGrid == "linear") {
    for (int n = 0; n < npt; n++) T[n] = s.T1 + n * d0;
  } else {
    // get_gauss_TGrid (CF/subrouts.cpp:4104-4140), alpha = 0.2
    const double dx = 10.0 / (npt - 1), alpha = 0.2;
    std::vector<double> tv(npt, 0.0);
    for (int n = 0; n < npt - 1; n++) {
      const double x = (n + 1) * dx;
      tv[n + 1] = tv[n] + d0 * (1.0 - exp(-alpha * pow(x, 2)));
    }
    const double fc = (s.T2 - s.T1) / tv[npt - 1];
    for (int n = 0; n < npt; n++) T[n] = fc * tv[n] + s.T1;
  }
  return T;
}

std::string t_tag(double T, const std::string &label) { return "_T(" + fixed(T, 5) + ")" + label; }

// Per-bin averages of CF/mc_hm_code.cpp:5376-5636: Values1 (energy,
// Binder cumulants, susceptibilities, order parameters, spin stiffness) and
// Values2 (S(q) peaks and correlation lengths at the five key points, scaled
// by 1/(measCnt2 Gsz)).  Entries with |v| <= 1e-12 (or NaN) are written as 0.
void bin_values(const double *mom, int64_t n, int N, int Gsz, double beta, bool ising, bool c3,
                bool qc, double v1[23], double v2[40], double v3[45]) {
  const double fc1 = 1.0 / (double)n;
  double a[HSMD_N_MOMENTS];
  for (int k = 0; k < HSMD_N_MOMENTS; k++) a[k] = fc1 * mom[k];
  const double E1 = a[0], E2 = a[1], F1 = a[2], F2 = a[3], F4 = a[4], T1 = a[5], T2 = a[6];
  const double Q1 = a[7], Q2 = a[8], Q4 = a[9], S1 = a[10], S2 = a[11], S4 = a[12];
  const double SH = a[13], SV = a[14], SD = a[15], X1 = a[16], X2 = a[17], X4 = a[18];
  const double iN = 1.0 / N;
  const double EDen1 = iN * E1, EDen2 = iN * iN * E2;
  const double fs = N * beta, fh = N * beta * beta;
  double FB, QB, SB, XB;
  if (ising) {
    FB = F2 == 0 ? 0 : 1.5 * (1.0 - F4 / (3.0 * F2 * F2));
    QB = Q2 == 0 ? 0 : 1.5 * (1.0 - Q4 / (3.0 * Q2 * Q2));
    SB = S2 == 0 ? 0 : 2.0 * (1.0 - S4 / (2.0 * S2 * S2));
  } else {
    FB = F2 == 0 ? 0 : 0.5 * (5.0 - 3.0 * F4 / (F2 * F2));
    QB = Q2 == 0 ? 0 : 0.5 * (5.0 - 3.0 * Q4 / (Q2 * Q2));
    SB = S2 == 0 ? 0 : 1.0 * (4.0 - 3.0 * S4 / (S2 * S2));
  }
  if (c3) XB = 1.5 * (X2 == 0 ? 0 : 1.0 - X4 / (2.0 * X2 * X2));
  else XB = 1.5 * (X2 == 0 ? 0 : 1.0 - X4 / (3.0 * X2 * X2));
  const double fs0 = N * beta;
  for (int k = 0; k < 45; k++) v3[k] = 0.0;
  if (qc) {
    // quasicrystal Binder forms (CF/mc_hm_code.cpp:5471-5495) and Values3
    // (StPrj / SqPks / N1fpar blocks, CF/mc_hm_code.cpp:5547-5640)
    QB = Q2 == 0 ? 0 : 1.0 - Q4 / (3.0 * Q2 * Q2);
    SB = S2 == 0 ? 0 : 1.5 - S4 / (S2 * S2);
    XB = 1.5 * (X2 == 0 ? 0 : 1.0 - X4 / (X2 * X2));
    for (int k = 0; k < 4; k++) {
      const double p1 = a[66 + k], p2 = a[70 + k], p4 = a[74 + k];
      const double s1 = a[81 + k], s2 = a[85 + k], s4 = a[89 + k];
      const double row1[5] = {p1, p2, p4, p2 > 0.0 ? 1.0 - p4 / (3.0 * p2 * p2) : 0.0, fs0 * (p2 - p1 * p1)};
      const double row2[5] = {s1, s2, s4, s2 > 0.0 ? 2.0 - s4 / (s2 * s2) : 0.0, fs0 * (s2 - s1 * s1)};
      for (int j = 0; j < 5; j++) {
        v3[5 * k + j] = row1[j];
        v3[20 + 5 * k + j] = row2[j];
      }
    }
    const double n1 = a[78], n2 = a[79], n4 = a[80];
    const double row3[5] = {n1, n2, n4, n2 == 0.0 ? 0.0 : 1.0 - n4 / (3.0 * n2 * n2), fs0 * (n2 - n1 * n1)};
    for (int j = 0; j < 5; j++) v3[40 + j] = row3[j];
    for (int k = 0; k < 45; k++)
      if (!(fabs(v3[k]) > 1.0e-12)) v3[k] = 0.0;
  }
  // spin stiffness, fac_RhoSt = -2 / (sqrt(3) N) (CF/mc_hm_code.cpp:3741, 5411-5415)
  const double fr = -2.0 / (sqrt(3.0) * N);
  const double RhoS1 = fr * (a[19] + a[20] + a[21]) / 3.0;
  const double RhoS2 = beta * fr * (a[22] + a[23] + a[24]) / 3.0;
  const double vals[23] = {EDen1, EDen2, fh * (EDen2 - EDen1 * EDen1),
                           QB, fs * (Q2 - Q1 * Q1), SB, fs * (S2 - S1 * S1),
                           XB, fs * (X2 - X1 * X1), FB, fs * (F2 - F1 * F1),
                           fs * (T2 - T1 * T1), Q1, S1, fabs(X1), F1, T1,
                           SH, SV, SD, RhoS1, RhoS2, RhoS1 + RhoS2};
  for (int k = 0; k < 23; k++) v1[k] = fabs(vals[k]) > 1.0e-12 ? vals[k] : 0.0;
  const double fc2 = 1.0 / mom[65];
  for (int k = 0; k < 40; k++) {
    const double v = (fc2 / Gsz) * mom[25 + k];
    v2[k] = fabs(v) > 1.0e-12 ? v : 0.0;
  }
}

// MC_Bin_w0SqField (CF/mc_hm_code.cpp:5752-5800): int32 count, then per q
// the complex (sum |S_a|) e^{i <arg S_a>} for a = x, y, z.
void write_w0_sq_field(const std::string &path, const double *f, int Nq, int64_t count) {
  std::ofstream out(path, std::ios::binary | std::ios::trunc);
  if (!out) fail(HSMD_ERR_IO, "cannot write " + path);
  const int32_t cnt = (int32_t)count;
  out.write(reinterpret_cast<const char *>(&cnt), sizeof(cnt));
  const double fc = 1.0 / (double)count;
  for (int q = 0; q < Nq; q++) {
    for (int d = 0; d < 3; d++) {
      const double amp = f[(size_t)q * 6 + d], ph = fc * f[(size_t)q * 6 + 3 + d];
      const double z[2] = {amp * cos(ph), amp * sin(ph)};
      out.write(reinterpret_cast<const char *>(z), sizeof(z));
    }
  }
}

}  // namespace

// Full MC run: thermalisation, binned measurement with replica exchange,
// spin-field recording and the per-bin averages (CF/mc_hm_code.cpp:139).
void run_mc(const std::string &dir, int npt, int precision, uint64_t seed, bool strict) {
  Settings s = read_settings(dir, strict);
  if (npt < 1) fail(HSMD_ERR_ARG, "replica count must be >= 1");
  const bool pc = npt > 1;
  if (pc && s.code[PT_ON] && npt % 2 != 0) fail(HSMD_ERR_GUARD, "parallel tempering needs an even number of replicas");
  LatHolder lh;
  make_lattice(s, strict, lh);
  int32_t info[16];
  check(hsmd_lattice_info(lh.p, info));
  const int N = info[0];
  const bool c3 = info[15] != 0;
  std::vector<int32_t> imp;
  if (s.imp > 0.0) {
    imp = impurity_map(N, s.imp, seed);
    check(hsmd_lattice_set_impurities(lh.p, imp.data()));
  }
  const bool qc = s.code[QCRYSTAL] != 0;
  if (qc) {
    // reference states of the projection order parameters
    // (CF/mc_hm_code.cpp:1747-1860; zeros with refState_Zero)
    std::vector<double> neel((size_t)N * 3, 0.0), strp((size_t)4 * N * 3, 0.0);
    if (!s.code[REFSTATE_ZERO]) {
      const std::string tag = "_AppxtSz(" + std::to_string(N) + ").bin";
      const char *names[5] = {"stateNeel0", "stateStrp1", "stateStrp2", "stateStrp3", "stateStrp4"};
      for (int r = 0; r < 5; r++) {
        const std::string f = s.dir + "/outBin/" + names[r] + tag;
        long long sz = 0;
        if (!file_size(f, sz) || sz != (long long)N * 24) fail(HSMD_ERR_IO, "Invalid data in: " + f);
        const std::vector<double> v = read_bin<double>(f, (size_t)3 * N);
        std::copy(v.begin(), v.end(), r == 0 ? neel.begin() : strp.begin() + (size_t)(r - 1) * N * 3);
      }
    } else if (s.IState == "inputf") {
      fail(HSMD_ERR_GUARD, "Input initial qct-state is invalid, reference states must be available");
    }
    check(hsmd_lattice_set_qc_states(lh.p, neel.data(), strp.data(), s.code[QCT_NEELPHASE] ? 1 : 0, nullptr));
  }
  const std::vector<double> T = temperature_grid(s, npt);
  hsmd_mc_spec sp;
  sp.model = s.code[ISI_MODEL] ? HSMD_ISING : HSMD_HEISENBERG;
  sp.precision = precision;
  sp.n_or = 5; sp.n_hb = 5; sp.n_or_boost = 2; sp.n_hb_boost = 8;
  sp.pt_cap50 = 1;
  sp.seed = seed;
  // initial states first: host-only, so input errors surface before any GPU work
  std::vector<double> spins((size_t)npt * N * 3);
  for (int r = 0; r < npt; r++) initial_state(s, s.L, N, c3, r, seed, &spins[(size_t)r * N * 3]);
  hsmd_mc *mc = nullptr;
  check(hsmd_mc_create(lh.p, npt, T.data(), &sp, &mc));
  std::unique_ptr<hsmd_mc, void (*)(hsmd_mc *)> guard(mc, hsmd_mc_destroy);
  check(hsmd_mc_set_spins(mc, spins.data()));
  const bool pt = pc && s.code[PT_ON];
  // thermalisation: exchange from sweep 101 on, pair parity = sweep % 2
  const int n0 = std::min(s.NTerm, 101);
  check(hsmd_mc_run(mc, n0, -1, 0, nullptr));
  if (s.NTerm > n0) check(hsmd_mc_run(mc, s.NTerm - n0, pt ? 2 : -1, 0, nullptr));
  // measurement: NBins bins of BinSz sweeps, pair parity = bin index % 2
  check(hsmd_mc_reset_counters(mc));
  // S(q) every dftDelay1 sweeps of a bin: 10, or BinSz for tabulated graphs
  // (CF/subrouts.cpp:2143-2153)
  const bool dft = s.code[WITH_DFT] != 0;
  const int Gsz = info[6], Nq = info[7];
  if (dft) check(hsmd_mc_set_sq(mc, s.code[QCRYSTAL] ? kBinSz : 10));
  std::vector<double> dummy(HSMD_N_MOMENTS * (size_t)npt);
  int64_t nm = 0;
  check(hsmd_mc_moments(mc, dummy.data(), &nm, 1));
  const int NBins = s.NMeas / kBinSz;
  std::vector<std::string> f1(npt), f2(npt), f3(npt), f0(npt);
  for (int k = 0; k < npt; k++) {
    const std::string tt = t_tag(T[k], s.label) + ".bin";
    f0[k] = dir + "/outBin/MC_Bin_SpinField" + tt;
    f1[k] = dir + "/outBin/MC_Bin_AvgValue1" + tt;
    f2[k] = dir + "/outBin/MC_Bin_AvgValue2" + tt;
    f3[k] = dir + "/outBin/MC_Bin_AvgValue3" + tt;
  }
  const bool rec = s.code[WITH_RECFIELD] != 0;
  if (rec && !imp.empty())
    for (int k = 0; k < npt; k++)
      if (T[k] <= kTempMax) write_bin(f0[k], imp.data(), imp.size(), true);
  std::vector<double> mom((size_t)npt * HSMD_N_MOMENTS);
  for (int b = 0; b < NBins; b++) {
    const int par = pt ? (b % 2) : -1;
    for (int c = 0; c < kBinSz; c += s.crSkip) {
      check(hsmd_mc_run(mc, s.crSkip, par, 1, nullptr));
      if (rec) {
        check(hsmd_mc_get_spins(mc, spins.data()));
        for (int k = 0; k < npt; k++)
          if (T[k] <= kTempMax) write_bin(f0[k], &spins[(size_t)k * N * 3], (size_t)N * 3, true);
      }
    }
    check(hsmd_mc_moments(mc, mom.data(), &nm, 1));
    for (int k = 0; k < npt; k++) {
      double v1[23], v2[40], v3[45];
      bin_values(&mom[(size_t)k * HSMD_N_MOMENTS], nm, N, Gsz, 1.0 / T[k], s.code[ISI_MODEL] != 0,
                 c3, qc, v1, v2, v3);
      write_bin(f1[k], v1, 23, true);
      write_bin(f2[k], v2, 40, true);
      write_bin(f3[k], v3, 45, true);
    }
  }
  if (qc) {
    std::vector<int32_t> hist((size_t)npt * HSMD_QC_HIST);
    check(hsmd_mc_qc_histograms(mc, hist.data(), 0));
    for (int k = 0; k < npt; k++)
      write_qc_histograms(dir, &hist[(size_t)k * HSMD_QC_HIST], t_tag(T[k], s.label));
  }
  if (dft) {
    std::vector<double> sqf((size_t)npt * Nq * 6);
    int64_t cnt = 0;
    check(hsmd_mc_sq_field(mc, sqf.data(), &cnt, 0));
    for (int k = 0; k < npt; k++) {
      const double *f = &sqf[(size_t)k * Nq * 6];
      write_w0_sq_field(dir + "/outBin/MC_Bin_w0SqField" + t_tag(T[k], s.label) + ".bin", f, Nq, cnt);
      // averaged S(q) (fc (sum |S|) e^{i <arg S>}) -> zone maps and path files
      std::vector<double> sq((size_t)Nq * 6);
      const double fc = 1.0 / (double)cnt;
      for (int q = 0; q < Nq; q++)
        for (int d = 0; d < 3; d++) {
          const double amp = fc * f[(size_t)q * 6 + d], ph = fc * f[(size_t)q * 6 + 3 + d];
          sq[((size_t)q * 3 + d) * 2] = amp * cos(ph);
          sq[((size_t)q * 3 + d) * 2 + 1] = amp * sin(ph);
        }
      write_mc_spectra(dir, lh.p->lat, sq.data(), t_tag(T[k], s.label) + ".dat");
    }
  }
  write_bin(dir + "/outBin/Temp_List" + s.label + ".bin", T.data(), T.size(), false);
  // statistics of the bin averages (CF/mc_hm_code.cpp:6430-7239)
  if (NBins >= 2) {
    std::vector<std::string> tags(npt);
    for (int k = 0; k < npt; k++) tags[k] = t_tag(T[k], s.label);
    const std::string info = s.outTag == "Htag" ? "  Extn. Magnetic Field"
                             : (s.outTag == "J2tag" ? "  Coupling strength J2" : "  Coupling strength J1");
    const double lv = s.outTag == "Htag" ? s.h : (s.outTag == "J2tag" ? s.J2 : s.J1);
    write_mc_stats(dir, s.label, lv, info, T, tags, (uint32_t)seed);
  }
}

// Full SMD run over every recorded sample at temperature TpVec[tpIndex]
// (CF/smd_code.cpp:141): RK4 + correction + DSSF, then the sample-averaged
// MC_Bin_SpecData file (CF/smd_code.cpp:3590-3630).
void run_smd(const std::string &dir, int precision, bool strict) {
  Settings s = read_settings(dir, strict);
  const int ntm = s.ntm;
  const double tmax = ntm * s.dtm, dwf = 2.0 * acos(-1.0) / tmax;
  if (dwf > kDwfMin && strict)
    fail(HSMD_ERR_GUARD, "Frequency discretization dwf = 2 pi / (ntm dtm) exceeds dwf_min = 0.10; increase ntm or dtm");
  LatHolder lh;
  make_lattice(s, strict, lh);
  int32_t info[16];
  check(hsmd_lattice_info(lh.p, info));
  const int N = info[0], Nq = info[7];
  const std::string tfile = dir + "/outBin/Temp_List" + s.label + ".bin";
  auto TpVec = read_bin<double>(tfile, (size_t)s.tpLstSz);
  const double Temp = TpVec[s.tpIndex];
  // samples file: packs of [N int32 impurity map if disorder] + NMeas/crSkip configurations
  const std::string sf = dir + "/outBin/MC_Bin_SpinField" + t_tag(Temp, s.label) + ".bin";
  long long sz;
  if (!file_size(sf, sz)) fail(HSMD_ERR_IO, "File not found: " + sf);
  const bool dis = s.imp > 0.0;
  const long long szConf = (long long)N * 24, szImap = (long long)N * 4;
  const long long packSize = s.NMeas / s.crSkip;
  const long long szPack = (dis ? szImap : 0) + packSize * szConf;
  long long nPacks, perPack;
  if (!dis && sz % szConf == 0) { nPacks = 1; perPack = sz / szConf; }
  else if (dis && sz % szPack == 0) { nPacks = sz / szPack; perPack = packSize; }
  else fail(HSMD_ERR_IO, "Error: input samples file has no valid data! Input: " + sf);
  std::ifstream in(sf, std::ios::binary);
  std::vector<hsmd_lattice *> lats;
  std::vector<std::unique_ptr<LatHolder>> packLats;
  std::vector<int32_t> sample_lat;
  std::vector<double> spins;
  for (long long p = 0; p < nPacks; p++) {
    if (dis) {
      std::vector<int32_t> imp(N);
      in.read(reinterpret_cast<char *>(imp.data()), szImap);
      packLats.emplace_back(new LatHolder());
      make_lattice(s, strict, *packLats.back());
      check(hsmd_lattice_set_impurities(packLats.back()->p, imp.data()));
      lats.push_back(packLats.back()->p);
    } else if (p == 0) {
      lats.push_back(lh.p);
    }
    const size_t off = spins.size();
    spins.resize(off + (size_t)perPack * N * 3);
    in.read(reinterpret_cast<char *>(&spins[off]), perPack * szConf);
    for (long long k = 0; k < perPack; k++) sample_lat.push_back((int32_t)(lats.size() - 1));
  }
  if (!in) fail(HSMD_ERR_IO, "Error: inconsistency found in the input samples file! File : " + sf);
  const int S = (int)sample_lat.size();
  // temporal window (CF/smd_code.cpp:2808-2895); "wread" reads outBin/winvec.bin
  std::vector<double> w(ntm, 1.0);
  if (s.tWForm == "wread") {
    w = read_bin<double>(dir + "/outBin/winvec.bin", (size_t)ntm);
    double mx = 0;
    for (double v : w) mx = std::max(mx, v);
    for (double &v : w) v /= mx;
  } else if (!temporal_window(s.tWForm, ntm, w.data())) {
    fail(HSMD_ERR_GUARD, "Invalid temporal window (please check inputs);");
  }
  hsmd_smd_spec sp;
  sp.precision = precision; sp.dt = s.dtm; sp.n_steps = ntm; sp.n_omega = ntm; sp.dssf = 1;
  sp.gauss_seidel = 0;
  hsmd_smd *smd = nullptr;
  std::vector<int32_t> sl(sample_lat.begin(), sample_lat.end());
  check(hsmd_smd_create(lats.data(), (int)lats.size(), sl.data(), S, &sp, w.data(), &smd));
  std::unique_ptr<hsmd_smd, void (*)(hsmd_smd *)> guard(smd, hsmd_smd_destroy);
  check(hsmd_smd_set_samples(smd, spins.data()));
  check(hsmd_smd_step(smd, ntm - 1));
  std::vector<double> acc((size_t)3 * Nq * ntm, 0.0), part(acc.size()), q(ntm, 0.0), qp(ntm);
  for (int li = 0; li < (int)lats.size(); li++) {
    check(hsmd_smd_dssf(smd, li, part.data(), qp.data()));
    for (size_t k = 0; k < acc.size(); k++) acc[k] += part[k];
    for (int n = 0; n < ntm; n++) q[n] += qp[n];
  }
  const double fc = 1.0 / S;
  std::vector<double> out;
  out.reserve((size_t)3 * Nq * ntm + ntm);
  for (int k = 0; k < Nq; k++)
    for (int d = 0; d < 3; d++)
      for (int n = 0; n < ntm; n++) out.push_back(fc * acc[((size_t)d * Nq + k) * ntm + n]);
  for (int n = 0; n < ntm; n++) out.push_back(fc * q[n]);
  write_bin(dir + "/outBin/MC_Bin_SpecData" + t_tag(Temp, s.label) + ".bin", out.data(), out.size(), false);
  // text outputs (quantum correction, path, w-slices, qSS), CF/smd_code.cpp:3628-4130
  if (s.code[WITH_DFT])
    write_smd_spectra(dir, lh.p->lat, out.data(), out.data() + (size_t)3 * Nq * ntm, ntm, s.dtm,
                      Temp, t_tag(Temp, s.label) + ".dat");
}

}  // namespace hsmd

extern "C" {

hsmd_status hsmd_run_mc(const char *workdir, int n_replicas, int precision, uint64_t seed,
                        int strict) {
  try {
    hsmd::run_mc(workdir ? workdir : ".", n_replicas, precision, seed, strict != 0);
    return HSMD_OK;
  } catch (const hsmd::Error &e) {
    hsmd::set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    hsmd::set_last_error(e.what());
    return HSMD_ERR_ARG;
  }
}

hsmd_status hsmd_smd_window(const char *name, int n, double *out) {
  try {
    if (!name || !out) throw hsmd::Error(HSMD_ERR_ARG, "null argument");
    if (n < 2) throw hsmd::Error(HSMD_ERR_ARG, "a window needs at least 2 points");
    if (!hsmd::temporal_window(name, n, out))
      throw hsmd::Error(HSMD_ERR_ARG, std::string("unknown temporal window ") + name);
    return HSMD_OK;
  } catch (const hsmd::Error &e) {
    hsmd::set_last_error(e.what());
    return e.code;
  }
}

hsmd_status hsmd_run_smd(const char *workdir, int precision, int strict) {
  try {
    hsmd::run_smd(workdir ? workdir : ".", precision, strict != 0);
    return HSMD_OK;
  } catch (const hsmd::Error &e) {
    hsmd::set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    hsmd::set_last_error(e.what());
    return HSMD_ERR_ARG;
  }
}

}  // extern "C"
