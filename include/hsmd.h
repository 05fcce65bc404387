/*
 * hsmd.h -- C ABI of the MI355X-native Monte Carlo sweep engine and
 * semiclassical spin-dynamics (SMD) engine for 2D Heisenberg / Ising models.
 *
 * The reference (rmpt-phys/Classical_MC_and_SMD_for_2D_Heisenberg_Models) has
 * no library boundary: its hot path is a set of functions over mutable globals
 * called from two monolithic `main`s.  Each entry point below names the
 * reference function or inline block it replaces (paths relative to the
 * reference root, CF = Codes_Folder).
 *
 * ABI rules
 *   - every function returns an hsmd_status (0 = success); on failure the
 *     message is kept per thread and read back with hsmd_last_error();
 *   - all buffers are caller-owned host memory; the library copies in and out
 *     and keeps device memory opaque;
 *   - spins cross the ABI as FP64 (x,y,z) triples in site order, the layout of
 *     the reference's binary sample files (CF/mc_hm_code.cpp:5167-5177);
 *   - no C++ exception crosses the ABI; handles are not thread-safe (one
 *     handle per host thread), but distinct handles are independent.
 */
#ifndef HSMD_H
#define HSMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int hsmd_status;
enum {
  HSMD_OK = 0,
  HSMD_ERR_ARG = 1,     /* invalid argument / shape */
  HSMD_ERR_GUARD = 2,   /* a reference guard rail rejected the input */
  HSMD_ERR_DEVICE = 3,  /* HIP runtime error */
  HSMD_ERR_IO = 4,      /* file missing / wrong size */
  HSMD_ERR_STATE = 5    /* call out of order (e.g. run before set_spins) */
};

/* Precision / batching forms of the MC and SMD engines. */
enum {
  HSMD_PLAIN_F64 = 0,    /* FP64, one instance per launch, global memory */
  HSMD_BATCHED_F32 = 1,  /* FP32 state, FP64 reductions, all instances in one launch */
  HSMD_BATCHED_F64 = 2   /* FP64 state, all instances in one launch */
};

enum { HSMD_HEISENBERG = 0, HSMD_ISING = 1 };

/* Coupling constants, one-to-one with Settings_MODEL.txt / Settings_Disorder.txt
 * (Set_Params.py:6-72; parsed at CF/subrouts.cpp:955-2186). */
typedef struct hsmd_couplings {
  double J1, J2, JX;      /* crystal-crystal couplings per shell */
  double J2_delta;        /* Lieb-lattice J2 asymmetry (requires geometry "lieb") */
  double lambda1, lambda2;/* XXZ factors: shell 1 uses lambda1, shells 2 and X use lambda2 */
  double h;               /* external field along z (dMag = 2) */
  double imp_ratio;       /* impurity ratio in [0,1); 0 disables disorder */
  double J1ic, J2ic, JXic;/* impurity-crystal couplings */
  double J1ii, J2ii, JXii;/* impurity-impurity couplings */
} hsmd_couplings;

typedef struct hsmd_lattice hsmd_lattice;
typedef struct hsmd_mc hsmd_mc;
typedef struct hsmd_smd hsmd_smd;

/* Copies the last error message of the calling thread (NUL-terminated). */
hsmd_status hsmd_last_error(char *buf, size_t len);

/* Number of visible HIP devices (0 when none); never fails. */
int hsmd_device_count(void);

/* ------------------------------------------------------------------ lattice */

/* Builds a periodic lattice with the reference's site numbering, neighbour
 * shells and in-shell slot order (replaces make_rvecList CF/subrouts.cpp:2714,
 * make_nborsTable CF/subrouts.cpp:2983, make_bondList CF/subrouts.cpp:3426).
 * geometry: "square", "triang", "hexagn", "kagome", "lieb".
 * pbc = 0 turns image neighbours into -1 holes (the reference's PBC_OFF).
 * strict = 1 applies the reference's guard rails (L even; L % 3 == 0 for
 * C3-symmetric geometries); strict = 0 is the extended mode. */
hsmd_status hsmd_lattice_generate(const char *geometry, int L, int pbc, int strict,
                                  const hsmd_couplings *c, hsmd_lattice **out);

/* Builds a lattice from a tabulated neighbour table: per site zn1 shell-1,
 * then zn2 shell-2, then znx shell-X int32 entries, -1 = missing (the
 * nborsList*.bin layout written at CF/subrouts.cpp:3294-3316, read at
 * CF/mc_hm_code.cpp:1273-1320).  rvec (n x 2 f64, may be NULL) is the
 * sitesList*.bin layout (CF/mc_hm_code.cpp:955-1020).  The direct-sum q
 * grid of get_qct_SqData (CF/subrouts.cpp:8460) is Lsz x Lsz with Lsz the even
 * integer part of sqrt(n) (CF/subrouts.cpp:1104-1112). */
hsmd_status hsmd_lattice_from_tables(int32_t n, int zn1, int zn2, int znx,
                                     const int32_t *nbors, const double *rvec,
                                     const hsmd_couplings *c, hsmd_lattice **out);
/* Seeded Poisson-disk point set in a periodic box [0, box)^2 with minimum
 * distance r_min (Bridson's algorithm); writes up to max_points points
 * (x, y f64) and their number.  Generates the irregular-graph inputs that the
 * reference takes as tabulated files (its QCrystal_Data converters); no
 * counterpart function in the reference. */
hsmd_status hsmd_poisson_disk(double box, double r_min, uint64_t seed, int32_t max_points,
                              double *points, int32_t *n_points);
/* Tabulated lattice from points in a periodic box: shell 1 = minimum-image
 * distance < r_shell[0], shell 2 < r_shell[1], shell X < r_shell[2]; slots by
 * distance; widths = largest counts (an extended nborsList layout beyond the
 * reference's fixed 8/3/8, CF/subrouts.cpp:1012).  Then as
 * hsmd_lattice_from_tables. */
hsmd_status hsmd_lattice_from_points(int32_t n, const double *rvec, double box,
                                     const double r_shell[3], const hsmd_couplings *c,
                                     hsmd_lattice **out);

/* Sets the impurity map (n int32 of 0/1; CF/subrouts.cpp:3321-3352 draws it). */
hsmd_status hsmd_lattice_set_impurities(hsmd_lattice *lat, const int32_t *imp);

/* Removes sites: every table entry pointing at a vacant site becomes -1 and
 * the vacant site loses all its own entries (a vacancy is a site whose
 * couplings are all zero). */
hsmd_status hsmd_lattice_set_vacancies(hsmd_lattice *lat, const int32_t *vacant);

/* info[16]: n, zn1, zn2, znx, n_bonds, n_colours, Gsz, Nq, nsubs, L,
 *           j2_on, jx_on, xxz_on, h_on, disorder, c3sym */
hsmd_status hsmd_lattice_info(const hsmd_lattice *lat, int32_t info[16]);
hsmd_status hsmd_lattice_tables(const hsmd_lattice *lat, int32_t *nbors, double *rvec);
hsmd_status hsmd_lattice_grid_map(const hsmd_lattice *lat, int32_t *grid_map);
/* bonds: n_bonds x 2 int32 pairs (i, j) in make_bondList order; shell: n_bonds int32 (1,2,3). */
hsmd_status hsmd_lattice_bonds(const hsmd_lattice *lat, int32_t *pairs, int32_t *shell);
/* colour of each site (0..n_colours-1) used by the GPU sweeps. */
hsmd_status hsmd_lattice_colours(const hsmd_lattice *lat, int32_t *colour);
/* Colour schedule of the GPU sweeps (no counterpart in the reference, which
 * visits sites in a shuffled order, get_shuffledList CF/subrouts.cpp:2666):
 * cap = 0 restores the default (first-fit in site order, rebalanced to
 * ceil(n / n_colours) sites per colour when that costs at most two more
 * colours); cap > 0 is first-fit with at most cap sites per colour.  Engines
 * created afterwards use it; hsmd_mc_autotune picks the cap. */
hsmd_status hsmd_lattice_set_colour_cap(hsmd_lattice *lat, int32_t cap);
/* Quasicrystal reference states for the projection order parameters:
 * neel[n x 3], stripes[4 x n x 3] (stateNeel0 / stateStrp1..4_AppxtSz(n).bin,
 * read at CF/mc_hm_code.cpp:1747-1860); the S(q) hotspots are located from
 * the Neel state (neel_phase = 1) or the first stripe state (qct_NeelPhase,
 * get_qct_SqPkLoc CF/subrouts.cpp:8531).  MC engines created afterwards
 * measure the quasicrystal observables.  peaks (nullable): 8 x 2 int32
 * (nx, ny) in units of the lattice's dq. */
hsmd_status hsmd_lattice_set_qc_states(hsmd_lattice *lat, const double *neel,
                                       const double *stripes, int neel_phase,
                                       int32_t *peaks);
void hsmd_lattice_destroy(hsmd_lattice *lat);

/* --------------------------------------------------- fields and energies */

/* Local fields of R spin configurations (R x n x 3 f64 in, R x n x 3 f64 out);
 * replaces get_localField (CF/subrouts.cpp:4465) evaluated at every site.
 * precision: HSMD_PLAIN_F64 / HSMD_BATCHED_F64 (FP64) or HSMD_BATCHED_F32. */
hsmd_status hsmd_local_fields(const hsmd_lattice *lat, int precision, int R,
                              const double *spins, double *fields);

/* Total energies of R configurations (R f64); replaces get_energyValue
 * (CF/subrouts.cpp:5181).  The reduction is FP64 in every precision form. */
hsmd_status hsmd_energies(const hsmd_lattice *lat, int precision, int R,
                          const double *spins, double *energies);

/* ---------------------------------------------------------------- MC engine */

typedef struct hsmd_mc_spec {
  int model;        /* HSMD_HEISENBERG or HSMD_ISING */
  int precision;    /* HSMD_PLAIN_F64 / HSMD_BATCHED_F32 / HSMD_BATCHED_F64 */
  int n_or, n_hb;   /* overrelaxation and heat-bath passes per sweep (reference: 5, 5) */
  int n_or_boost, n_hb_boost; /* passes in the sweep after an accepted swap (reference: 2, 8) */
  int pt_cap50;     /* 1: reference 50 % acceptance-ratio cap (CF/mc_hm_code.cpp:3099) */
  uint64_t seed;
} hsmd_mc_spec;

/* Result of hsmd_mc_autotune. */
typedef struct hsmd_mc_tuning {
  int32_t colour_cap;   /* for hsmd_lattice_set_colour_cap (0 = default schedule) */
  int32_t threads;      /* for hsmd_mc_set_threads */
  double ms_per_sweep;  /* measured with the chosen schedule and shape */
  double default_ms;    /* measured with the default schedule and 1024 threads */
  int32_t n_candidates; /* (schedule, shape) pairs timed; 0 when cached */
  int32_t cached;       /* 1: read from the tuning cache (HSMD_TUNE_CACHE) */
} hsmd_mc_tuning;

/* R replicas, temps[R] per temperature slot (the reference's rank order). */
hsmd_status hsmd_mc_create(const hsmd_lattice *lat, int R, const double *temps,
                           const hsmd_mc_spec *spec, hsmd_mc **out);
/* Spins of the configuration at each temperature slot (R x n x 3 f64).
 * Ising spins are (+-1, 0, 0) as in the reference (CF/subrouts.cpp:4695-4725). */
hsmd_status hsmd_mc_set_spins(hsmd_mc *mc, const double *spins);
/* Threads per replica workgroup of the batched forms (multiple of 64 in
 * [64, 1024]; default 1024).  No counterpart in the reference (one MPI rank
 * per replica). */
hsmd_status hsmd_mc_set_threads(hsmd_mc *mc, int threads);
/* Autotuner of the batched sweep (no counterpart in the reference): times
 * n_sweeps sweeps of R replicas on lat for every distinct colour schedule
 * (first-fit colourings with caps of k x 64..1024 sites per colour, the
 * default, no cap) and workgroup shape (256, 512, 1024 threads) on the
 * device, and returns the fastest.  lat is not modified; apply the result
 * with hsmd_lattice_set_colour_cap and hsmd_mc_set_threads.  With
 * HSMD_TUNE_CACHE naming a file, results are cached there per lattice
 * signature, R, model, precision, passes and device. */
hsmd_status hsmd_mc_autotune(const hsmd_lattice *lat, int R, const double *temps,
                             const hsmd_mc_spec *spec, int n_sweeps, hsmd_mc_tuning *out);
hsmd_status hsmd_mc_get_spins(hsmd_mc *mc, double *spins);
/* Name of the sweep kernel the engine launches in its current form (kernel,
 * template arguments, threads per workgroup; " wave-ordered" when the fast
 * kernel orders its colour classes per wave rather than by a workgroup
 * barrier), at most len bytes with the terminator.  Introspection with no
 * counterpart in the reference. */
hsmd_status hsmd_mc_sweep_kernel(hsmd_mc *mc, char *name, int len);

/* Runs n_sweeps sweeps of every replica; replaces the sweep loops
 * CF/mc_hm_code.cpp:2778-2962 (thermalisation) and 4836-4945 (measurement)
 * and the MPI replica exchange CF/mc_hm_code.cpp:3015-3130 / 5253-5365.
 * pt_parity: -1 = no exchange; 2 = alternate pair parity with the engine's
 * sweep counter (thermalisation); 0 or 1 = fixed parity (measurement uses
 * the bin index).  measure = 1 accumulates the per-slot observable moments.
 * energies (nullable): n_sweeps x R f64, the energy at each slot after each
 * sweep, before the exchange. */
hsmd_status hsmd_mc_run(hsmd_mc *mc, int n_sweeps, int pt_parity, int measure,
                        double *energies);
/* Energies (get_energyValue) of n_configs configurations held in the
 * caller's device memory, spins_dev = n_configs x n x 3 f64 (site order),
 * written to the caller's device buffer energies_dev (n_configs f64), with
 * the bond list of mc's lattice; launched on `stream` (a hipStream_t; NULL =
 * the null stream) without waiting for it.  No counterpart in the reference:
 * the variant of hsmd_energies for callers that cannot allocate device memory
 * (code under graph capture, tight serving loops).  It allocates nothing: the
 * partial sums of its reduction live in a fixed per-device scratch array of
 * 2^18 FP64 values, so n_configs x ceil(n_bonds / 4096) must not exceed
 * 262144 (HSMD_ERR_ARG otherwise).  Calls that overlap on one GPU share that
 * array and race. */
hsmd_status hsmd_mc_config_energies_noalloc(hsmd_mc *mc, int n_configs, const double *spins_dev,
                                            double *energies_dev, void *stream);
/* Energies per temperature slot of the current configurations. */
hsmd_status hsmd_mc_energies(hsmd_mc *mc, double *energies);
/* attempts[R], accepts[R]: the reference's per-rank ptCnt / ptExc counters. */
hsmd_status hsmd_mc_pt_counters(hsmd_mc *mc, int64_t *attempts, int64_t *accepts);
/* Resets the exchange counters and the sweep counter (CF/mc_hm_code.cpp:4767). */
hsmd_status hsmd_mc_reset_counters(hsmd_mc *mc);
/* Observable moments accumulated per slot since the last reset:
 * out[R x HSMD_N_MOMENTS], n_meas = number of measured sweeps.
 * Moment order (CF/mc_hm_code.cpp:4951-5245):
 *   [0..18]  E, E^2, F1, F2, F4, T1, T2, Q1, Q2, Q4, S1, S2, S4, SH, SV, SD,
 *            X1, X2, X4 (X = C3 Psi or Z2 Ising parameter);
 *   [19..24] spin stiffness (get_spinStiff, CF/subrouts.cpp:6519-6620):
 *            sum rhoS1vec[0..2], sum rhoS2vec[0..2] (the squared C-sums);
 *   [25..44] S(q) peak values SQ1 at the five key points, xy, yz, zz, tt
 *            (get_corrLen's first output, stored by the reference as
 *            xyCLen..ttCLen, Values2[0..19]);
 *   [45..64] correlation lengths (get_corrLen's second output, xySVal..ttSVal,
 *            Values2[20..39]);
 *   [65]     number of S(q) measurements (measCnt2);
 *   [66..92] quasicrystal lattices with reference states
 *            (hsmd_lattice_set_qc_states; CF/mc_hm_code.cpp:4973-5160):
 *            StPrjVc1[4], StPrjVc2[4], StPrjVc4[4], N1fpar 1/2/4,
 *            SqPks4P1[4], SqPks4P2[4], SqPks4P4[4].  On such lattices
 *            Q, S, X of [7..18] are the projection forms (|Neel projection|,
 *            |stripe projections|, |stripe vector| of get_qctStripeMag).
 * [25..65] accumulate only when S(q) measurement is on (hsmd_mc_set_sq). */
#define HSMD_N_MOMENTS 93
/* Per-slot quasicrystal histograms (int32): Neel projection [41], stripe
 * projections [41][4], stripe vector [41 (imag)][41 (real)]
 * (NeelHist / StrpHist / SMagVecHist, CF/mc_hm_code.cpp:4496-4745). */
#define HSMD_QC_HIST 1886
hsmd_status hsmd_mc_moments(hsmd_mc *mc, double *out, int64_t *n_meas, int reset);
/* The quasicrystal histograms per slot, out[R x HSMD_QC_HIST] (counts since
 * the last reset; they are not cleared by hsmd_mc_moments). */
hsmd_status hsmd_mc_qc_histograms(hsmd_mc *mc, int32_t *out, int reset);
/* Static S(q) measurement during measured sweeps: at every `every`-th
 * measured sweep counted from the last hsmd_mc_moments reset, starting with
 * the first (the reference's nbin % dftDelay1 == 0 within a bin,
 * CF/mc_hm_code.cpp:5185-5245); 0 turns it off.
 * Periodic generated lattices use the grid DFT of get_SqData_iMKL
 * (CF/subrouts.cpp:8330-8443), tabulated or open ones the direct sums of
 * get_qct_SqData (CF/subrouts.cpp:8460-8525). */
hsmd_status hsmd_mc_set_sq(hsmd_mc *mc, int every);
/* Accumulated S(q) per slot, out[R x Nq x 6] = sum |S_x|, |S_y|, |S_z|,
 * arg S_x, arg S_y, arg S_z (the reference's SqField,
 * CF/mc_hm_code.cpp:5205-5213); count = number of S(q) measurements. */
hsmd_status hsmd_mc_sq_field(hsmd_mc *mc, double *out, int64_t *count, int reset);
/* Static S(q) of n_configs configurations (n_configs x n x 3 f64, the
 * reference's site order): out[n_configs x 3 x Nq x 2] (re, im), the output
 * of get_SqData_iMKL / get_qct_SqData (same choice as hsmd_mc_set_sq). */
hsmd_status hsmd_static_sq(const hsmd_lattice *lat, int n_configs, const double *spins,
                           double *out);
/* Slot -> configuration map (R int32). */
hsmd_status hsmd_mc_slot_map(hsmd_mc *mc, int32_t *config_of_slot);
void hsmd_mc_destroy(hsmd_mc *mc);

/* -------------------------------------------------------------- SMD engine */

typedef struct hsmd_smd_spec {
  int precision;    /* HSMD_PLAIN_F64 / HSMD_BATCHED_F32 / HSMD_BATCHED_F64 */
  double dt;        /* time step dtm */
  int n_steps;      /* ntm: number of time points, including t = 0 */
  int n_omega;      /* frequency bins kept, n < n_omega (n_omega = n_steps keeps all) */
  int dssf;         /* 1: accumulate the DSSF; 0: dynamics only */
  int gauss_seidel; /* 1: in-place sequential energy correction (reference order,
                       CPU only); 0: Jacobi (GPU) */
} hsmd_smd_spec;

/* S samples; sample s evolves on lattices[sample_lattice[s]]. window[n_steps]
 * is the temporal window (nullable = "nowin").  With dssf set, the engine
 * buffers trajectory chunks on the device and contracts each over time when
 * it fills: up to 8192 points (FP32) or 256 (FP64), fewer when the chunk
 * would not fit in 90 % of the free device memory; the environment variable
 * HSMD_SMD_MAX_CHUNK (a power of two >= 256) caps the FP32 chunk. */
hsmd_status hsmd_smd_create(const hsmd_lattice *const *lattices, int n_lattices,
                            const int32_t *sample_lattice, int S,
                            const hsmd_smd_spec *spec, const double *window,
                            hsmd_smd **out);
/* Result of hsmd_smd_autotune. */
typedef struct hsmd_smd_tuning {
  int32_t form;              /* for hsmd_smd_set_shape: 1, 2 or 3 */
  int32_t gemm;              /* for hsmd_smd_set_shape: 0..6 */
  double ms_per_step;        /* RK4 step of the whole batch, chosen form */
  double default_ms_per_step;/* same, form chosen at creation */
  double gemm_ms;            /* contraction of one trajectory chunk of up to 2048 points,
                                chosen tile (-1: no DSSF); for the FFT form (6), whose
                                chunk costs a whole transform, its cost over n_steps
                                points in chunks of up to 8192, scaled to the chunk */
  double default_gemm_ms;
  int32_t n_candidates;      /* shapes timed; 0 when cached */
  int32_t cached;            /* 1: read from the tuning cache (HSMD_TUNE_CACHE) */
} hsmd_smd_tuning;

/* Launch shape of the batched forms (no counterpart in the reference, which
 * runs one sample per rank).  form: 0 = keep, 1 = LDS-resident with the
 * batch's largest slot class, 2 = LDS-resident with per-workgroup slot
 * classes (FP32 batches of two classes), 3 = global-memory form.  gemm:
 * 0 = 128 x 128 DSSF contraction tile, 2 = 128 x 256, 3 = 128 x 128 with
 * two LDS stages (all three FP32 with 2 n_omega a multiple of 256),
 * 4 = 128 x 512 (FP32 with 2 n_omega a multiple of 512), 5 = the 128 x 512
 * tile pipelined: twiddles in registers, trajectory through a three-stage
 * LDS ring (same preconditions and sums as 4), 6 = the time DFT as a pruned
 * FFT (FP32 with n_steps a multiple of 512 and n_omega <= 512),
 * 1 = 64 x 64.
 * Errors when the shape does
 * not apply to the batch. */
hsmd_status hsmd_smd_set_shape(hsmd_smd *smd, int form, int gemm);
/* Name of the RK4 kernel hsmd_smd_step launches in the engine's current form
 * (kernel, template arguments, threads per workgroup; ", wave-ordered on k of
 * n lattices" when the LDS-resident form orders its RK stages per wave
 * rather than by workgroup barriers), at most len bytes with the terminator.
 * Introspection with no counterpart in the reference. */
hsmd_status hsmd_smd_step_kernel(hsmd_smd *smd, char *name, int len);
/* Autotuner of the batched SMD step (no counterpart in the reference): on a
 * scratch engine of the same batch, times n_steps RK4 steps per applicable
 * form and one trajectory-chunk contraction per tile, and returns the
 * fastest of each; apply with hsmd_smd_set_shape.  Cached like
 * hsmd_mc_autotune. */
hsmd_status hsmd_smd_autotune(const hsmd_lattice *const *lattices, int n_lattices,
                              const int32_t *sample_lattice, int S,
                              const hsmd_smd_spec *spec, const double *window,
                              int n_steps, hsmd_smd_tuning *out);
/* Spins of every sample, concatenated in sample order (sum_s n_s x 3 f64). */
hsmd_status hsmd_smd_set_samples(hsmd_smd *smd, const double *spins);
/* MC -> SMD hand-off on the device: sample s starts from the configuration
 * at temperature slot slots[s] of mc (every sample must be on the MC
 * engine's lattice).  Replaces the round trip through MC_Bin_SpinField
 * (written at CF/mc_hm_code.cpp:5167-5177, read at CF/smd_code.cpp:3079-3130)
 * for in-process pipelines; same effect as hsmd_smd_set_samples with those
 * spins. */
hsmd_status hsmd_smd_set_samples_from_mc(hsmd_smd *smd, hsmd_mc *mc, const int32_t *slots);
hsmd_status hsmd_smd_get_samples(hsmd_smd *smd, double *spins);
/* Advances every sample n_steps RK4 steps (each with its energy correction);
 * replaces the time loop of get_dynSpectrum (CF/subrouts.cpp:11734-11927).
 * The first call records t = 0.  The sum of calls must not exceed
 * spec.n_steps - 1 when the DSSF is on. */
hsmd_status hsmd_smd_step(hsmd_smd *smd, int n_steps);
/* Appends n_points time points of given configurations to the trajectory of
 * every sample instead of integrating them (spins: n_points blocks, each laid
 * out as hsmd_smd_set_samples); the first call supplies t = 0.  The points
 * take the same trajectory chunks, time contraction and spatial transform as
 * integrated ones, so hsmd_smd_dssf then returns the DSSF of an externally
 * computed trajectory (the reference's own SMD output, another integrator),
 * and qss its autocorrelation against the first point.  Replaces the time
 * transform of get_dynSpectrum applied to a stored history
 * (CF/subrouts.cpp:12062-12105).  An engine that was fed is not stepped. */
hsmd_status hsmd_smd_feed(hsmd_smd *smd, int n_points, const double *spins);
/* Sample-summed DSSF of lattice li: out[3][Nq][n_omega] f64 =
 * sum_s |X_s(q, w)|^2 with X the windowed, 1/sqrt(Nq T)-normalised space-time
 * DFT (CF/subrouts.cpp:11906-11925, 12062-12105); for tabulated graphs the
 * spatial transform is get_qct_SqData's direct sum on its L x L q grid,
 * normalised 1/sqrt(N T) (CF/subrouts.cpp:8460-8525), and qss[n_steps] the
 * sample-summed autocorrelation (1/N) sum_i S_i(0).S_i(t).  Available after
 * n_steps - 1 steps. Divide by the sample count for the reference's averages. */
hsmd_status hsmd_smd_dssf(hsmd_smd *smd, int li, double *out, double *qss);
/* Sample-summed DSSF of lattice li at nq arbitrary wavevectors q[nq x 2]
 * (a q-path), by direct spatial sums: out[3][nq][n_omega] f64 =
 * sum_s |sum_r e^{-i q.r} Y_s(r, w)|^2 / (N T), Y the windowed time DFT of the
 * spins; the transform get_qct_SqData applies to tabulated graphs
 * (CF/subrouts.cpp:8460-8525), at the caller's q.  Vacant sites are left out.
 * No counterpart call in the reference, which fixes q to its grid. */
hsmd_status hsmd_smd_dssf_q(hsmd_smd *smd, int li, int nq, const double *q, double *out);
/* Energies of every sample (S f64). */
hsmd_status hsmd_smd_energies(hsmd_smd *smd, double *energies);
void hsmd_smd_destroy(hsmd_smd *smd);

/* ----------------------------------------------------------- drop-in runs */

/* Full MC run in workdir (replaces main of CF/mc_hm_code.cpp:139, launched
 * there as `mpirun --np n_replicas ./xMC_run`): reads the five Settings_*.txt
 * files, builds the lattice (writing outBin/nborsList.bin), thermalises,
 * measures in bins of 1000 sweeps with replica exchange, and appends
 * outBin/MC_Bin_AvgValue{1,2,3}_T(..)<label>.bin and
 * outBin/MC_Bin_SpinField_T(..)<label>.bin, then writes
 * outBin/Temp_List<label>.bin.  strict = 1 keeps every reference guard rail. */
hsmd_status hsmd_run_mc(const char *workdir, int n_replicas, int precision, uint64_t seed,
                        int strict);

/* Full SMD run in workdir (replaces main of CF/smd_code.cpp:141): evolves
 * every sample of outBin/MC_Bin_SpinField_T(TpVec[tpIndex])<label>.bin and
 * writes the sample-averaged outBin/MC_Bin_SpecData_T(..)<label>.bin. */
hsmd_status hsmd_run_smd(const char *workdir, int precision, int strict);

/* Temporal window of the DSSF time transform on n >= 2 points, as the SMD
 * run builds it from Settings' tWForm (CF/smd_code.cpp:2808-2895): "nowin",
 * "gauss", "lancz", "sincw", "hann1", "hann2", "hann3", "black", "ntall",
 * "flatw" ("wread", read from outBin/winvec.bin, is the run's own).  The
 * result is the window argument of hsmd_smd_create. */
hsmd_status hsmd_smd_window(const char *name, int n, double *out);

/* --------------------------------------------------------- statistics */

/* Mean and standard deviation over bins of n_values bin averages
 * (data: n_bins x n_values, bin-major) -- get_statVectors
 * (CF/aux_code.cpp:1050-1187).  bootstrap = 1: deviation of the means of
 * 1000 resamples with replacement drawn with MT19937 seeded by `seed` and
 * GSL's uniform_int rule (the reference's USE_GSL branch); 0: the sample
 * standard deviation. */
hsmd_status hsmd_bin_statistics(const double *data, int n_bins, int n_values, int bootstrap,
                                uint32_t seed, double *mean, double *sdev);

/* --------------------------------------------------------- self checks */

/* First n outputs of the MT19937 generator the bootstrap uses. */
hsmd_status hsmd_debug_mt19937(uint32_t seed, int n, uint32_t *out);

/* Writes n Philox4x32-10 blocks for counters (site = i, step, inst, stream)
 * computed on the device (4n uint32), for the bit-exact stream check. */
hsmd_status hsmd_debug_philox(uint64_t seed, uint32_t step, uint32_t inst,
                              uint32_t stream, int n, uint32_t *out);
/* Launches n colour sweeps of every replica back to back (no measurement,
 * no exchange) and waits for them: times the sweep kernel alone.  Advances
 * the sweep counter like hsmd_mc_run. */
hsmd_status hsmd_debug_mc_sweeps(hsmd_mc *mc, int n);
/* Device det_exp_neg on n inputs, for the bit-exact acceptance check. */
hsmd_status hsmd_debug_exp(int n, const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif /* HSMD_H */
