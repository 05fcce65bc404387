// Statistics of the bin averages and the MC_Stat_Results / MC_Collect files
// written at the end of an MC run (CF/mc_hm_code.cpp:6430-7239).
//
// Per temperature and per value: the mean over bins and its standard
// deviation by bootstrap resampling (get_statVectors with USE_GSL,
// CF/aux_code.cpp:1050-1187: 1000 resamples with replacement, the standard
// deviation of their means).  The reference draws the resample indices with
// GSL's default generator; GSL is not a dependency here, so its published
// algorithms are restated:
//   * gsl_rng_mt19937 (Matsumoto and Nishimura 1998, the 2002 seeding
//     mt[0] = s (4357 for s = 0), mt[i] = 1812433253 (mt[i-1] ^ mt[i-1] >> 30) + i),
//   * gsl_rng_uniform_int(n): k = x / floor((2^32 - 1) / n), rejecting k >= n,
//   * gsl_stats_mean / gsl_stats_sd: running mean and variance in extended
//     precision, sd = sqrt(var n / (n - 1)).
// The simple (non-bootstrap) branch is the sample standard deviation.
#include <math.h>
#include <stdio.h>
#include <sys/stat.h>

#include <fstream>
#include <string>
#include <vector>

#include "hsmd_internal.h"

namespace hsmd {

struct Mt19937 {
  uint32_t mt[624];
  int mti = 624;
  explicit Mt19937(uint32_t s) {
    if (s == 0) s = 4357;
    mt[0] = s;
    for (int i = 1; i < 624; i++) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
  }
  uint32_t next() {
    if (mti >= 624) {
      for (int k = 0; k < 624; k++) {
        const uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
        mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
      }
      mti = 0;
    }
    uint32_t y = mt[mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  uint32_t uniform_int(uint32_t n) {
    const uint32_t scale = 0xffffffffu / n;
    uint32_t k;
    do {
      k = next() / scale;
    } while (k >= n);
    return k;
  }
};

static double running_mean(const double *x, size_t n) {
  long double m = 0;
  for (size_t i = 0; i < n; i++) m += (x[i] - m) / (long double)(i + 1);
  return (double)m;
}

static double running_sd(const double *x, size_t n) {
  const double m = running_mean(x, n);
  long double v = 0;
  for (size_t i = 0; i < n; i++) {
    const long double d = x[i] - m;
    v += (d * d - v) / (long double)(i + 1);
  }
  return sqrt((double)v * ((double)n / (double)(n - 1)));
}

// data: n_bins x n_values (bin-major); mean, sdev: n_values.
void bin_statistics(const double *data, int n_bins, int n_values, bool bootstrap, uint32_t seed,
                    double *mean, double *sdev) {
  if (n_bins < 2 || n_values < 1) throw Error(HSMD_ERR_ARG, "need at least two bins and one value");
  std::vector<double> col(n_bins), res(n_bins), means(1000);
  for (int v = 0; v < n_values; v++) {
    for (int b = 0; b < n_bins; b++) col[b] = data[(size_t)b * n_values + v];
    if (bootstrap) {
      mean[v] = running_mean(col.data(), n_bins);
      // a fresh generator per value, as the reference allocates one per call
      Mt19937 rng(seed);
      for (int r = 0; r < 1000; r++) {
        for (int j = 0; j < n_bins; j++) res[j] = col[rng.uniform_int((uint32_t)n_bins)];
        means[r] = running_mean(res.data(), n_bins);
      }
      sdev[v] = running_sd(means.data(), 1000);
    } else {
      double m = 0.0;
      for (int b = 0; b < n_bins; b++) m += col[b];
      m *= 1.0 / n_bins;
      double s = 0.0;
      for (int b = 0; b < n_bins; b++) s += (col[b] - m) * (col[b] - m);
      mean[v] = m;
      sdev[v] = sqrt(s / (n_bins - 1.0));
    }
  }
}

void debug_mt19937(uint32_t seed, int n, uint32_t *out) {
  Mt19937 rng(seed);
  for (int i = 0; i < n; i++) out[i] = rng.next();
}

namespace {

const char *kNames1[23] = {"EnrgDens (p1)", "EnrgDens (p2)", "Sys spec-heat", "Q-Binder4thCM",
                           "Q-mag suscept", "S-Binder4thCM", "S-mag suscept", "X-Binder4thCM",
                           "X-mag suscept", "F-Binder4thCM", "F-mag suscept", "T-mag suscept",
                           "Q-magnet (p1)", "S-magnet (p1)", "X-magnet (p1)", "F-magnet (p1)",
                           "T-magnet (p1)", "S-magnet (HS)", "S-magnet (VS)", "S-magnet (DS)",
                           "Spin stiff C1", "Spin stiff C2", "Sys stiffness"};

std::string sci(double v) {
  char b[64];
  snprintf(b, sizeof b, "%+22.12e", v);
  return b;
}

std::string two(int n) {
  char b[16];
  snprintf(b, sizeof b, "%02d", n);
  return b;
}

std::vector<std::string> names(int which) {
  std::vector<std::string> v;
  const char *tag[4] = {"(xy)", "(yz)", "(zz)", "(tt)"};
  const char *ptag[4] = {"(p0)", "(p1)", "(p2)", "(p3)"};
  if (which == 1) {
    for (int i = 0; i < 23; i++) v.push_back(kNames1[i]);
  } else if (which == 2) {
    const char *k[5] = {"KP", "MX", "MY", "GP", "ZP"};
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 5; j++) v.push_back(std::string("SCLen ") + k[j] + " " + tag[i]);
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 5; j++) v.push_back(std::string("SpVal ") + k[j] + " " + tag[i]);
  } else {
    const char *a[5] = {"StPrjVc1 ", "StPrjVc2 ", "StPrjVc4 ", "StPrjBin ", "StPrjChi "};
    const char *b[5] = {"SpPks4P1 ", "SpPks4P2 ", "SpPks4P4 ", "SqPksBin ", "SqPksChi "};
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 5; j++) v.push_back(std::string(a[j]) + ptag[i]);
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 5; j++) v.push_back(std::string(b[j]) + ptag[i]);
    const char *c[5] = {"N1fpar1  ", "N1fpar2  ", "N1fpar4  ", "N1fBinC4 ", "chiN1fpar"};
    for (int j = 0; j < 5; j++) v.push_back(std::string(c[j]) + "    ");
  }
  return v;
}

// collect_data (CF/aux_code.cpp:1309-1356): append to `out` the lines of
// `in` that start with `key`, preceded by `in`'s first line when `out` is new.
void collect(const std::string &out, const std::string &in, const std::string &key) {
  std::ifstream probe(out);
  bool header = !probe.good();
  probe.close();
  std::ofstream o(out, std::ios::app);
  std::ifstream i(in);
  std::string line;
  while (std::getline(i, line)) {
    if (header) {
      o << line << "\n";
      header = false;
    }
    if (line.compare(0, key.size(), key) == 0) o << line << "\n";
  }
}

}  // namespace

// The MC_Stat_Results{1,2,3}<label>.dat tables (temperature, label value,
// then mean and deviation of every bin value) and their per-temperature
// copies MC_Collect/MC_Stat_FixTemp{1,2,3}(T).dat, appended across runs.
// tags[k] is the "_T(..)<label>" tag of temperature k; seed_k = seed + k
// replaces the reference's per-rank random GSL seed.  The reference builds
// the FixTemp names from a string stream it never clears, so from the second
// temperature on its names run the temperatures together; here each file is
// named by its own temperature.
void write_mc_stats(const std::string &dir, const std::string &label, double label_value,
                    const std::string &info, const std::vector<double> &T,
                    const std::vector<std::string> &tags, uint32_t seed) {
  const int M[3] = {23, 40, 45};
  const std::string base = dir + "/outData/";
  mkdir(base.c_str(), 0755);
  mkdir((base + "MC_Collect").c_str(), 0755);
  const std::string X3 = "   ";
  for (int w = 0; w < 3; w++) {
    const std::string path = base + "MC_Stat_Results" + std::to_string(w + 1) + label + ".dat";
    std::ofstream out(path, std::ios::trunc);
    if (!out) throw Error(HSMD_ERR_IO, "cannot write " + path);
    const auto nm = names(w + 1);
    out << "# Temperature (kB = 1)" << X3 << info << X3;
    for (int n = 0; n < M[w]; n++) {
      const int m = 2 * (n + 1);
      out << nm[n] << " (MV) c" << two(m + 1) << X3 << nm[n] << " (DV) c" << two(m + 2) << X3;
    }
    out << "\n";
    for (size_t k = 0; k < T.size(); k++) {
      const std::string f = dir + "/outBin/MC_Bin_AvgValue" + std::to_string(w + 1) + tags[k] + ".bin";
      std::ifstream in(f, std::ios::binary | std::ios::ate);
      if (!in) throw Error(HSMD_ERR_IO, "could not open input binary file " + f);
      const long long sz = (long long)in.tellg();
      const int nb = (int)(sz / (long long)(M[w] * sizeof(double)));
      if (nb < 2 || sz != (long long)nb * M[w] * (long long)sizeof(double))
        throw Error(HSMD_ERR_IO, "bin file of an unexpected size: " + f);
      in.seekg(0);
      std::vector<double> d((size_t)nb * M[w]);
      in.read(reinterpret_cast<char *>(d.data()), sz);
      std::vector<double> mean(M[w]), sd(M[w]);
      bin_statistics(d.data(), nb, M[w], true, seed + (uint32_t)k, mean.data(), sd.data());
      out << sci(T[k]) << X3 << sci(label_value) << X3;
      for (int n = 0; n < M[w]; n++) out << sci(mean[n]) << X3 << sci(sd[n]) << X3;
      out << "\n";
    }
    out.close();
    for (size_t k = 0; k < T.size(); k++) {
      char t[64];
      snprintf(t, sizeof t, "%.5f", T[k]);
      collect(base + "MC_Collect/MC_Stat_FixTemp" + std::to_string(w + 1) + "(" + t + ").dat", path,
              sci(T[k]));
    }
  }
}

}  // namespace hsmd
