// Autotuner support: signatures of lattices and devices, and the optional
// on-disk cache of tuning results.
//
// The cache is a text file named by the environment variable
// HSMD_TUNE_CACHE, one entry per line: "<key> <v0> <v1> ...".  Later entries
// for a key replace earlier ones.  Without the variable nothing is read or
// written and every tuning call measures.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "hsmd_internal.h"

namespace hsmd {

// FNV-1a, 64 bit
uint64_t fnv1a(const void *data, size_t n, uint64_t h) {
  const unsigned char *p = (const unsigned char *)data;
  for (size_t i = 0; i < n; i++) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

template <typename T>
static uint64_t mix_vec(uint64_t h, const std::vector<T> &v) {
  const size_t n = v.size();
  h = fnv1a(&n, sizeof n, h);
  return v.empty() ? h : fnv1a(v.data(), v.size() * sizeof(T), h);
}

// Everything the sweep and step kernels see of a lattice: the tables, the
// impurity and vacancy maps and the couplings (not the colouring, which is
// what the tuner chooses).
uint64_t lattice_signature(const Lattice &L) {
  uint64_t h = fnv1a(L.geom.data(), L.geom.size());
  const int dims[6] = {L.N, L.zn1, L.zn2, L.znx, L.L, L.pbc ? 1 : 0};
  h = fnv1a(dims, sizeof dims, h);
  h = mix_vec(h, L.nbors);
  h = mix_vec(h, L.imp);
  h = mix_vec(h, L.vacant);
  h = fnv1a(&L.c, sizeof L.c, h);
  return h;
}

std::string device_tag() {
  int dev = 0;
  hipDeviceProp_t p{};
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess)
    return "nodevice";
  std::string arch = p.gcnArchName;
  for (char &ch : arch)
    if (ch == ' ' || ch == ':') ch = '_';
  return arch + "_cu" + std::to_string(p.multiProcessorCount);
}

bool tune_cache_get(const std::string &key, std::vector<double> &vals) {
  const char *path = getenv("HSMD_TUNE_CACHE");
  if (!path || !*path) return false;
  std::ifstream f(path);
  if (!f) return false;
  std::string line;
  bool found = false;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string k;
    if (!(is >> k) || k != key) continue;
    std::vector<double> v;
    double x;
    while (is >> x) v.push_back(x);
    vals = v;
    found = true;
  }
  return found;
}

void tune_cache_put(const std::string &key, const std::vector<double> &vals) {
  const char *path = getenv("HSMD_TUNE_CACHE");
  if (!path || !*path) return;
  FILE *f = fopen(path, "a");
  if (!f) throw Error(HSMD_ERR_IO, std::string("cannot append to tuning cache ") + path);
  fprintf(f, "%s", key.c_str());
  for (double v : vals) fprintf(f, " %.17g", v);
  fprintf(f, "\n");
  fclose(f);
}

}  // namespace hsmd
