// Internal host-side types of the hsmd library (not part of the C ABI).
#pragma once
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hsmd.h"

namespace hsmd {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

// Maximum number of coupled slots per site the kernels accept.
constexpr int ZMAX = 32;
// Maximum number of colour classes.
constexpr int CMAX = 64;
// Maximum number of bond classes (distinct (Jxy, Jz) pairs).
constexpr int BCMAX = 16;

// Sublattice labellings used by the order parameters
// (CF/subrouts.cpp:5382-5800).  -1 = site not counted.
struct Labels {
  bool present = false;
  int n_stag = 2;                    // 2 (C2) or 3 (C3) sublattices
  std::vector<int8_t> stag, hs, vs, ds;  // original site order
  bool z2 = false;                   // square lattice: Ising-Z2 plaquette parameter
  std::vector<int32_t> z2nb;         // 3 x N original labels: slot1[0], slot1[1], slot2[0]
  bool c3 = false;                   // C3 Psi parameter (uses stag labels, z component)
};

struct Lattice {
  // --- reference description
  std::string geom;          // "square", ..., or "tables"
  int L = 0;                 // Lsz
  int nsubs = 1, gridFac = 1;
  int N = 0, Gsz = 0, Nq = 0;
  int zn1 = 0, zn2 = 0, znx = 0;
  bool tabulated = false;    // qcrystal-style input
  bool pbc = true;           // periodic boundaries (generated lattices)
  std::vector<int32_t> nbors;    // N x (zn1+zn2+znx), -1 = missing
  std::vector<double> rvec;      // N x 2
  std::vector<int32_t> gridMap;  // Gsz*Gsz -> site or -1 (periodic lattices)
  std::vector<int32_t> imp;      // N, 0/1
  std::vector<int32_t> vacant;   // N, 0/1
  hsmd_couplings c{};
  bool j2_on = false, jx_on = false, xxz_on = false, h_on = false;
  bool disorder = false, j2ab_on = false, c3sym = false;
  Labels labels;
  // colour schedule: 0 = default (first-fit, rebalanced when cheap);
  // > 0 = first-fit with at most col_cap sites per colour (autotuner)
  int col_cap = 0;
  // quasicrystal reference states (qc.cpp): 5 x N x 3 in original site
  // order (Neel, stripes 1-4) and the eight S(q) hotspots (nx, ny) in units
  // of sq_dq(); empty when not set
  std::vector<double> qc_ref;
  int qc_nx[8] = {0, 0, 0, 0, 0, 0, 0, 0}, qc_ny[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool qc_on() const { return !qc_ref.empty(); }

  // --- derived (rebuild())
  std::vector<int32_t> bi, bj, bshell;  // reference bond list (make_bondList order)
  std::vector<int32_t> colour;          // per original site
  int n_colours = 0;
  std::vector<int32_t> perm;            // new label -> original
  std::vector<int32_t> inv;             // original -> new label
  std::vector<int32_t> col_off;         // n_colours + 1
  int Z = 0;                            // active slots per site
  std::vector<int32_t> slot_shell;      // Z: 1, 2, 3
  std::vector<int32_t> slot_k;          // Z: slot index inside its shell
  bool general = false;                 // per-slot couplings needed
  double shellJ[3] = {0, 0, 0};         // uniform couplings per shell
  double shellLam[3] = {1, 1, 1};       // z-factor per shell
  // relabelled GPU tables
  std::vector<int32_t> g_nb;            // Z x N, new labels, N = zero sentinel
  std::vector<double> g_J;              // Z x N (general mode)
  std::vector<int32_t> g_bi, g_bj;      // bonds in new labels
  std::vector<uint8_t> g_bc;            // bond class
  std::vector<double> bc_Jxy, bc_Jz;    // per class
  std::vector<int32_t> g_perm;          // new -> original (copy of perm)

  void rebuild();
  // True when the bond-list energy equals 1/2 sum_i S_i.h_i - 1/2 h sum_i S_z,i
  // for every configuration (symmetric tables, each pair in one shell, and
  // field couplings equal to bond couplings); checked on a test configuration.
  bool field_energy_ok() const;
  double bondJ(int shell, int i, int j) const;   // J of a bond (energy)
  double stiffJ(int shell, int i, int j) const;  // J of a bond in the spin stiffness
  // static S(q): direct sums (tabulated or open lattices) or the grid DFT
  bool sq_direct() const { return tabulated || !pbc; }
  double sq_dq() const;   // momentum step dq (CF/mc_hm_code.cpp:1486)
  double slotJ(int shell, int k, int i, int nb) const;  // J of a field slot
  double lam(int shell) const { return shell == 1 ? (xxz_on ? c.lambda1 : 1.0)
                                                  : (xxz_on ? c.lambda2 : 1.0); }
};

// lattice.cpp
Lattice *lattice_generate(const std::string &geom, int L, bool pbc, bool strict,
                          const hsmd_couplings &c);
Lattice *lattice_from_tables(int n, int zn1, int zn2, int znx, const int32_t *nb,
                             const double *rvec, const hsmd_couplings &c);
void derive_flags(Lattice &lat);

// MC -> SMD hand-off on the device: where an MC engine keeps its
// configurations (mc.hip mc_handoff, read by smd.hip)
struct McHandoff {
  const void *spins;          // per configuration [3][N+1] (int8: [N+1])
  int type;                   // 0 int8 (Ising), 1 FP32, 2 FP64
  int R;
  const int32_t *cfg_of_slot; // device, R
  const Lattice *lat;         // labels: lat->perm (label -> site)
};

// stats.cpp: bin statistics and the MC_Stat_Results / MC_Collect files
void bin_statistics(const double *data, int n_bins, int n_values, bool bootstrap, uint32_t seed,
                    double *mean, double *sdev);
void debug_mt19937(uint32_t seed, int n, uint32_t *out);
void write_mc_stats(const std::string &dir, const std::string &label, double label_value,
                    const std::string &info, const std::vector<double> &T,
                    const std::vector<std::string> &tags, uint32_t seed);

// qc.cpp: quasicrystal reference states and hotspots
void lattice_set_qc(Lattice &L, const double *neel, const double *strp, bool neel_phase);

// tune.cpp: signature hashing and the optional on-disk tuning cache
// (path in the environment variable HSMD_TUNE_CACHE; no caching when unset)
uint64_t fnv1a(const void *data, size_t n, uint64_t h = 0xcbf29ce484222325ull);
uint64_t lattice_signature(const Lattice &L);
bool tune_cache_get(const std::string &key, std::vector<double> &vals);
void tune_cache_put(const std::string &key, const std::vector<double> &vals);
std::string device_tag();

}  // namespace hsmd

struct hsmd_lattice {
  hsmd::Lattice lat;
};
