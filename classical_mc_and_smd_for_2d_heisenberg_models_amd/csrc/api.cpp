// C-ABI layer of the hsmd library: argument checks, exception-to-status
// translation and per-thread error messages.  Every entry point is declared
// in include/hsmd.h with the reference function it replaces.
#include <hip/hip_runtime.h>
#include <string.h>

#include <memory>
#include <string>

#include "hsmd_internal.h"

namespace hsmd {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

struct MCEngine;
struct SMDEngine;
MCEngine *mc_create(const Lattice &L, int R, const double *temps, const hsmd_mc_spec &sp);
void mc_set_spins(MCEngine &e, const double *spins);
void mc_get_spins(MCEngine &e, double *spins);
void mc_run(MCEngine &e, int n_sweeps, int pt_parity, int measure, double *etrace);
void mc_energies(MCEngine &e, double *E);
void mc_refresh_energies(MCEngine &e);
void mc_pt_counters(MCEngine &e, int64_t *att, int64_t *acc);
void mc_reset_counters(MCEngine &e);
void mc_moments(MCEngine &e, double *out, int64_t *n, int reset);
void mc_slot_map(MCEngine &e, int32_t *m);
void mc_qc_histograms(MCEngine &e, int32_t *out, int reset);
void mc_set_sq(MCEngine &e, int every);
void mc_sq_field(MCEngine &e, double *out, int64_t *count, int reset);
void static_sq(const Lattice &L, int n_configs, const double *spins, double *out);
int poisson_disk(double box, double rmin, uint64_t seed, int max_n, double *out);
Lattice *lattice_from_points(int n, const double *rvec, double box, const double r[3],
                             const hsmd_couplings &c);
void mc_destroy(MCEngine *e);
void mc_config_energies_noalloc(const MCEngine &e, int R, const double *spins_dev,
                                double *energies_dev, hipStream_t st);
void mc_launch_sweeps(MCEngine &e, int n, hipStream_t st);
void mc_set_threads(MCEngine &e, int threads);
std::string mc_sweep_kernel(const MCEngine &e);
void mc_autotune(const Lattice &L, int R, const double *temps, const hsmd_mc_spec &sp, int n_sweeps,
                 hsmd_mc_tuning *out);
void fields_energies(const Lattice &L, int precision, int R, const double *spins, double *fields,
                     double *energies);
void debug_philox(uint64_t seed, uint32_t step, uint32_t inst, uint32_t stream, int n, uint32_t *out);
void debug_exp(int n, const double *x, double *y);

SMDEngine *smd_create(const Lattice *const *lats, int nl, const int32_t *sample_lattice, int S,
                      const hsmd_smd_spec &sp, const double *window, int max_chunk = 0);
void smd_set_samples(SMDEngine &e, const double *spins);
void smd_get_samples(SMDEngine &e, double *spins);
void smd_step(SMDEngine &e, int n);
void smd_feed(SMDEngine &e, int n, const double *spins);
void smd_dssf(SMDEngine &e, int li, double *out, double *qss);
void smd_dssf_q(SMDEngine &e, int li, int nq, const double *q, double *out);
void smd_energies(SMDEngine &e, double *E);
void smd_destroy(SMDEngine *e);
void smd_set_shape(SMDEngine &e, int form, int gemm);
std::string smd_step_kernel(const SMDEngine &e);
McHandoff mc_handoff(MCEngine &e);
void smd_set_samples_from_mc(SMDEngine &e, const McHandoff &h, const int32_t *slots);
void smd_autotune(const Lattice *const *lats, int nl, const int32_t *sample_lattice, int S,
                  const hsmd_smd_spec &sp, const double *window, int n_steps, hsmd_smd_tuning *out);

}  // namespace hsmd

using namespace hsmd;

struct hsmd_mc {
  MCEngine *e;
};
struct hsmd_smd {
  SMDEngine *e;
};

template <typename F>
static hsmd_status guarded(F &&f) {
  try {
    f();
    return HSMD_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return HSMD_ERR_ARG;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return HSMD_ERR_ARG;
  } catch (...) {
    set_last_error("unknown error");
    return HSMD_ERR_ARG;
  }
}

#define REQUIRE(cond, msg) \
  do { if (!(cond)) throw Error(HSMD_ERR_ARG, msg); } while (0)

extern "C" {

hsmd_status hsmd_last_error(char *buf, size_t len) {
  if (!buf || len == 0) return HSMD_ERR_ARG;
  const size_t n = std::min(len - 1, g_last_error.size());
  memcpy(buf, g_last_error.data(), n);
  buf[n] = 0;
  return HSMD_OK;
}

int hsmd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

hsmd_status hsmd_lattice_generate(const char *geometry, int L, int pbc, int strict,
                                  const hsmd_couplings *c, hsmd_lattice **out) {
  return guarded([&] {
    REQUIRE(geometry && c && out, "null argument");
    Lattice *lat = lattice_generate(geometry, L, pbc != 0, strict != 0, *c);
    hsmd_lattice *h = new hsmd_lattice();
    h->lat = std::move(*lat);
    delete lat;
    *out = h;
  });
}

hsmd_status hsmd_lattice_from_tables(int32_t n, int zn1, int zn2, int znx, const int32_t *nbors,
                                     const double *rvec, const hsmd_couplings *c,
                                     hsmd_lattice **out) {
  return guarded([&] {
    REQUIRE(nbors && c && out, "null argument");
    Lattice *lat = lattice_from_tables(n, zn1, zn2, znx, nbors, rvec, *c);
    hsmd_lattice *h = new hsmd_lattice();
    h->lat = std::move(*lat);
    delete lat;
    *out = h;
  });
}

hsmd_status hsmd_lattice_set_impurities(hsmd_lattice *lat, const int32_t *imp) {
  return guarded([&] {
    REQUIRE(lat && imp, "null argument");
    for (int i = 0; i < lat->lat.N; i++) REQUIRE(imp[i] == 0 || imp[i] == 1, "impurity map must be 0/1");
    lat->lat.imp.assign(imp, imp + lat->lat.N);
    lat->lat.rebuild();
  });
}

hsmd_status hsmd_lattice_set_vacancies(hsmd_lattice *lat, const int32_t *vac) {
  return guarded([&] {
    REQUIRE(lat && vac, "null argument");
    for (int i = 0; i < lat->lat.N; i++) REQUIRE(vac[i] == 0 || vac[i] == 1, "vacancy map must be 0/1");
    lat->lat.vacant.assign(vac, vac + lat->lat.N);
    lat->lat.rebuild();
  });
}

hsmd_status hsmd_lattice_set_colour_cap(hsmd_lattice *lat, int32_t cap) {
  return guarded([&] {
    REQUIRE(lat && cap >= 0, "bad argument");
    const int old = lat->lat.col_cap;
    lat->lat.col_cap = cap;
    try {
      lat->lat.rebuild();
    } catch (...) {
      lat->lat.col_cap = old;
      lat->lat.rebuild();
      throw;
    }
  });
}

hsmd_status hsmd_lattice_set_qc_states(hsmd_lattice *lat, const double *neel, const double *stripes,
                                       int neel_phase, int32_t *peaks) {
  return guarded([&] {
    REQUIRE(lat && neel && stripes, "null argument");
    lattice_set_qc(lat->lat, neel, stripes, neel_phase != 0);
    if (peaks)
      for (int i = 0; i < 8; i++) {
        peaks[2 * i] = lat->lat.qc_nx[i];
        peaks[2 * i + 1] = lat->lat.qc_ny[i];
      }
  });
}

hsmd_status hsmd_poisson_disk(double box, double r_min, uint64_t seed, int32_t max_points,
                              double *points, int32_t *n_points) {
  return guarded([&] {
    REQUIRE(points && n_points, "null argument");
    *n_points = poisson_disk(box, r_min, seed, max_points, points);
  });
}

hsmd_status hsmd_lattice_from_points(int32_t n, const double *rvec, double box,
                                     const double r_shell[3], const hsmd_couplings *c,
                                     hsmd_lattice **out) {
  return guarded([&] {
    REQUIRE(rvec && r_shell && c && out, "null argument");
    *out = nullptr;
    Lattice *L = lattice_from_points(n, rvec, box, r_shell, *c);
    hsmd_lattice *h = new hsmd_lattice();
    h->lat = std::move(*L);
    delete L;
    *out = h;
  });
}

hsmd_status hsmd_lattice_info(const hsmd_lattice *lat, int32_t info[16]) {
  return guarded([&] {
    REQUIRE(lat && info, "null argument");
    const Lattice &L = lat->lat;
    int32_t v[16] = {L.N, L.zn1, L.zn2, L.znx, (int32_t)L.bi.size(), L.n_colours, L.Gsz, L.Nq,
                     L.nsubs, L.L, L.j2_on, L.jx_on, L.xxz_on, L.h_on, L.disorder, L.c3sym};
    memcpy(info, v, sizeof(v));
  });
}

hsmd_status hsmd_lattice_tables(const hsmd_lattice *lat, int32_t *nbors, double *rvec) {
  return guarded([&] {
    REQUIRE(lat, "null argument");
    const Lattice &L = lat->lat;
    if (nbors) memcpy(nbors, L.nbors.data(), L.nbors.size() * sizeof(int32_t));
    if (rvec) {
      REQUIRE(!L.rvec.empty(), "lattice has no site coordinates");
      memcpy(rvec, L.rvec.data(), L.rvec.size() * sizeof(double));
    }
  });
}

hsmd_status hsmd_lattice_grid_map(const hsmd_lattice *lat, int32_t *gm) {
  return guarded([&] {
    REQUIRE(lat && gm, "null argument");
    REQUIRE(!lat->lat.gridMap.empty(), "lattice has no DFT grid");
    memcpy(gm, lat->lat.gridMap.data(), lat->lat.gridMap.size() * sizeof(int32_t));
  });
}

hsmd_status hsmd_lattice_bonds(const hsmd_lattice *lat, int32_t *pairs, int32_t *shell) {
  return guarded([&] {
    REQUIRE(lat, "null argument");
    const Lattice &L = lat->lat;
    for (size_t b = 0; b < L.bi.size(); b++) {
      if (pairs) { pairs[2 * b] = L.bi[b]; pairs[2 * b + 1] = L.bj[b]; }
      if (shell) shell[b] = L.bshell[b];
    }
  });
}

hsmd_status hsmd_lattice_colours(const hsmd_lattice *lat, int32_t *colour) {
  return guarded([&] {
    REQUIRE(lat && colour, "null argument");
    memcpy(colour, lat->lat.colour.data(), lat->lat.colour.size() * sizeof(int32_t));
  });
}

void hsmd_lattice_destroy(hsmd_lattice *lat) { delete lat; }

hsmd_status hsmd_local_fields(const hsmd_lattice *lat, int precision, int R, const double *spins,
                              double *fields) {
  return guarded([&] {
    REQUIRE(lat && spins && fields, "null argument");
    fields_energies(lat->lat, precision, R, spins, fields, nullptr);
  });
}

hsmd_status hsmd_energies(const hsmd_lattice *lat, int precision, int R, const double *spins,
                          double *energies) {
  return guarded([&] {
    REQUIRE(lat && spins && energies, "null argument");
    fields_energies(lat->lat, precision, R, spins, nullptr, energies);
  });
}

hsmd_status hsmd_mc_create(const hsmd_lattice *lat, int R, const double *temps,
                           const hsmd_mc_spec *spec, hsmd_mc **out) {
  return guarded([&] {
    REQUIRE(lat && temps && spec && out, "null argument");
    MCEngine *e = mc_create(lat->lat, R, temps, *spec);
    hsmd_mc *h = new hsmd_mc();
    h->e = e;
    *out = h;
  });
}

hsmd_status hsmd_mc_set_spins(hsmd_mc *mc, const double *spins) {
  return guarded([&] {
    REQUIRE(mc && spins, "null argument");
    mc_set_spins(*mc->e, spins);
    mc_refresh_energies(*mc->e);
  });
}

hsmd_status hsmd_mc_set_threads(hsmd_mc *mc, int threads) {
  return guarded([&] {
    REQUIRE(mc, "null argument");
    mc_set_threads(*mc->e, threads);
  });
}

hsmd_status hsmd_mc_sweep_kernel(hsmd_mc *mc, char *name, int len) {
  return guarded([&] {
    REQUIRE(mc && name && len > 0, "null argument");
    const std::string k = mc_sweep_kernel(*mc->e);
    snprintf(name, (size_t)len, "%s", k.c_str());
  });
}

hsmd_status hsmd_mc_autotune(const hsmd_lattice *lat, int R, const double *temps,
                             const hsmd_mc_spec *spec, int n_sweeps, hsmd_mc_tuning *out) {
  return guarded([&] {
    REQUIRE(lat && temps && spec && out, "null argument");
    mc_autotune(lat->lat, R, temps, *spec, n_sweeps, out);
  });
}

hsmd_status hsmd_mc_qc_histograms(hsmd_mc *mc, int32_t *out, int reset) {
  return guarded([&] {
    REQUIRE(mc && out, "null argument");
    mc_qc_histograms(*mc->e, out, reset);
  });
}

hsmd_status hsmd_mc_get_spins(hsmd_mc *mc, double *spins) {
  return guarded([&] {
    REQUIRE(mc && spins, "null argument");
    mc_get_spins(*mc->e, spins);
  });
}

hsmd_status hsmd_mc_run(hsmd_mc *mc, int n_sweeps, int pt_parity, int measure, double *energies) {
  return guarded([&] {
    REQUIRE(mc, "null argument");
    mc_run(*mc->e, n_sweeps, pt_parity, measure, energies);
  });
}

hsmd_status hsmd_mc_config_energies_noalloc(hsmd_mc *mc, int n_configs, const double *spins_dev,
                                            double *energies_dev, void *stream) {
  return guarded([&] {
    REQUIRE(mc && spins_dev && energies_dev, "null argument");
    mc_config_energies_noalloc(*mc->e, n_configs, spins_dev, energies_dev, (hipStream_t)stream);
  });
}

hsmd_status hsmd_mc_energies(hsmd_mc *mc, double *energies) {
  return guarded([&] {
    REQUIRE(mc && energies, "null argument");
    mc_energies(*mc->e, energies);
  });
}

hsmd_status hsmd_mc_pt_counters(hsmd_mc *mc, int64_t *attempts, int64_t *accepts) {
  return guarded([&] {
    REQUIRE(mc && attempts && accepts, "null argument");
    mc_pt_counters(*mc->e, attempts, accepts);
  });
}

hsmd_status hsmd_mc_reset_counters(hsmd_mc *mc) {
  return guarded([&] {
    REQUIRE(mc, "null argument");
    mc_reset_counters(*mc->e);
  });
}

hsmd_status hsmd_mc_moments(hsmd_mc *mc, double *out, int64_t *n_meas, int reset) {
  return guarded([&] {
    REQUIRE(mc && out, "null argument");
    mc_moments(*mc->e, out, n_meas, reset);
  });
}

hsmd_status hsmd_mc_set_sq(hsmd_mc *mc, int every) {
  return guarded([&] {
    REQUIRE(mc, "null argument");
    mc_set_sq(*mc->e, every);
  });
}

hsmd_status hsmd_mc_sq_field(hsmd_mc *mc, double *out, int64_t *count, int reset) {
  return guarded([&] {
    REQUIRE(mc, "null argument");
    mc_sq_field(*mc->e, out, count, reset);
  });
}

hsmd_status hsmd_static_sq(const hsmd_lattice *lat, int n_configs, const double *spins,
                           double *out) {
  return guarded([&] {
    REQUIRE(lat && spins && out, "null argument");
    static_sq(lat->lat, n_configs, spins, out);
  });
}

hsmd_status hsmd_mc_slot_map(hsmd_mc *mc, int32_t *config_of_slot) {
  return guarded([&] {
    REQUIRE(mc && config_of_slot, "null argument");
    mc_slot_map(*mc->e, config_of_slot);
  });
}

void hsmd_mc_destroy(hsmd_mc *mc) {
  if (!mc) return;
  mc_destroy(mc->e);
  delete mc;
}

hsmd_status hsmd_smd_set_samples_from_mc(hsmd_smd *smd, hsmd_mc *mc, const int32_t *slots) {
  return guarded([&] {
    REQUIRE(smd && mc && slots, "null argument");
    smd_set_samples_from_mc(*smd->e, mc_handoff(*mc->e), slots);
  });
}

hsmd_status hsmd_smd_step_kernel(hsmd_smd *smd, char *name, int len) {
  return guarded([&] {
    REQUIRE(smd && name && len > 0, "null argument");
    const std::string k = smd_step_kernel(*smd->e);
    snprintf(name, (size_t)len, "%s", k.c_str());
  });
}

hsmd_status hsmd_smd_set_shape(hsmd_smd *smd, int form, int gemm) {
  return guarded([&] {
    REQUIRE(smd, "null argument");
    smd_set_shape(*smd->e, form, gemm);
  });
}

hsmd_status hsmd_smd_autotune(const hsmd_lattice *const *lattices, int n_lattices,
                              const int32_t *sample_lattice, int S, const hsmd_smd_spec *spec,
                              const double *window, int n_steps, hsmd_smd_tuning *out) {
  return guarded([&] {
    REQUIRE(lattices && sample_lattice && spec && out && n_lattices > 0, "null argument");
    std::vector<const Lattice *> ls(n_lattices);
    for (int i = 0; i < n_lattices; i++) {
      REQUIRE(lattices[i], "null lattice");
      ls[i] = &lattices[i]->lat;
    }
    smd_autotune(ls.data(), n_lattices, sample_lattice, S, *spec, window, n_steps, out);
  });
}

hsmd_status hsmd_smd_create(const hsmd_lattice *const *lattices, int n_lattices,
                            const int32_t *sample_lattice, int S, const hsmd_smd_spec *spec,
                            const double *window, hsmd_smd **out) {
  return guarded([&] {
    REQUIRE(lattices && sample_lattice && spec && out && n_lattices > 0, "null argument");
    std::vector<const Lattice *> ls(n_lattices);
    for (int i = 0; i < n_lattices; i++) {
      REQUIRE(lattices[i], "null lattice");
      ls[i] = &lattices[i]->lat;
    }
    SMDEngine *e = smd_create(ls.data(), n_lattices, sample_lattice, S, *spec, window);
    hsmd_smd *h = new hsmd_smd();
    h->e = e;
    *out = h;
  });
}

hsmd_status hsmd_smd_set_samples(hsmd_smd *smd, const double *spins) {
  return guarded([&] {
    REQUIRE(smd && spins, "null argument");
    smd_set_samples(*smd->e, spins);
  });
}

hsmd_status hsmd_smd_get_samples(hsmd_smd *smd, double *spins) {
  return guarded([&] {
    REQUIRE(smd && spins, "null argument");
    smd_get_samples(*smd->e, spins);
  });
}

hsmd_status hsmd_smd_step(hsmd_smd *smd, int n_steps) {
  return guarded([&] {
    REQUIRE(smd, "null argument");
    smd_step(*smd->e, n_steps);
  });
}

hsmd_status hsmd_smd_feed(hsmd_smd *smd, int n_points, const double *spins) {
  return guarded([&] {
    REQUIRE(smd && (spins || n_points == 0), "null argument");
    smd_feed(*smd->e, n_points, spins);
  });
}

hsmd_status hsmd_smd_dssf_q(hsmd_smd *smd, int li, int nq, const double *q, double *out) {
  return guarded([&] {
    REQUIRE(smd && q && out, "null argument");
    smd_dssf_q(*smd->e, li, nq, q, out);
  });
}

hsmd_status hsmd_smd_dssf(hsmd_smd *smd, int li, double *out, double *qss) {
  return guarded([&] {
    REQUIRE(smd && out, "null argument");
    smd_dssf(*smd->e, li, out, qss);
  });
}

hsmd_status hsmd_smd_energies(hsmd_smd *smd, double *energies) {
  return guarded([&] {
    REQUIRE(smd && energies, "null argument");
    smd_energies(*smd->e, energies);
  });
}

void hsmd_smd_destroy(hsmd_smd *smd) {
  if (!smd) return;
  smd_destroy(smd->e);
  delete smd;
}

hsmd_status hsmd_debug_philox(uint64_t seed, uint32_t step, uint32_t inst, uint32_t stream, int n,
                              uint32_t *out) {
  return guarded([&] {
    REQUIRE(out && n >= 0, "bad argument");
    debug_philox(seed, step, inst, stream, n, out);
  });
}

hsmd_status hsmd_debug_mc_sweeps(hsmd_mc *mc, int n) {
  return guarded([&] {
    REQUIRE(mc && n >= 0, "bad argument");
    mc_launch_sweeps(*mc->e, n, 0);
    const hipError_t er = hipDeviceSynchronize();
    if (er != hipSuccess) throw Error(HSMD_ERR_DEVICE, std::string("sweeps: ") + hipGetErrorString(er));
  });
}

hsmd_status hsmd_bin_statistics(const double *data, int n_bins, int n_values, int bootstrap,
                                uint32_t seed, double *mean, double *sdev) {
  return guarded([&] {
    REQUIRE(data && mean && sdev, "null argument");
    bin_statistics(data, n_bins, n_values, bootstrap != 0, seed, mean, sdev);
  });
}

hsmd_status hsmd_debug_mt19937(uint32_t seed, int n, uint32_t *out) {
  return guarded([&] {
    REQUIRE(out && n >= 0, "bad argument");
    debug_mt19937(seed, n, out);
  });
}

hsmd_status hsmd_debug_exp(int n, const double *x, double *y) {
  return guarded([&] {
    REQUIRE(x && y && n >= 0, "bad argument");
    debug_exp(n, x, y);
  });
}

}  // extern "C"
