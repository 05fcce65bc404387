"""Benchmark of the MC sweep engine and the SMD engine on one MI355X.

Headline workload (value): batched parallel tempering, precision form --
triangular L=96 (N=9216), J1=1, J2=0.12, JX=0.05, h=0.3, XXZ lambda=0.9,
256 replicas geometric in T in [0.02, 2.0], FP32 spins and fields, FP64
energies and exchange acceptance, one sweep = 5 overrelaxation + 1 heat-bath
passes of every replica followed by the on-device replica exchange.  A step
is one such sweep of all replicas; value = spin-updates/s.

Also reported on the same line:
  roofline       -- the sweep kernel: VALU instructions per launch (committed
                    rocprofv3 summary) over the live launch time against the
                    SIMDs' issue peak, with the HBM view (algorithmic HBM bytes
                    of the LDS-resident design, measured HBM bytes) beside it;
  batched_f64 / plain_f64 / ising -- the same workload with FP64 state in one
                    launch, in the plain one-replica-per-launch FP64 form, and
                    as the Ising variant;
  config1        -- configuration 1 (square L=24, 16 replicas, 5e3 + 5e3
                    sweeps) in the plain FP64, batched FP64 and FP32 forms;
  smd            -- configuration 4: 512 samples over square L=64, kagome
                    L=32, hexagonal L=40 and square L=48 with 10 % vacancies,
                    FP32 RK4 (dt=0.005) with FP64 energy correction and the
                    DSSF time contraction (512 bins) included, in
                    site-RK4-steps/s; the RK4 kernel's and the contraction's
                    rooflines; the FP64 form of the batch; the DSSF error
                    against FP64; configuration 2 (32 samples, square L=24)
                    in the plain FP64, batched FP64 and FP32 forms;
  graph          -- configuration 5: about 12,000 seeded Poisson-disk points
                    (r_min = 1) in a periodic box, shells J1=1 (r<1.15),
                    J2=0.4 (<1.45), JX=0.1 (<1.75), 5 % vacancies; MC with
                    128 FP32 replicas, SMD with 256 samples over 4096 RK4
                    steps (dt=0.01) and the DSSF along a 200-point q-path;
  cpu_baseline   -- the oracle's reference-order FP64 chains (one thread per
                    replica, like one MPI rank each) on bounded samples of
                    configurations 3 and 1, and its reference-order SMD on
                    configuration 2 (one thread per sample).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import classical_mc_and_smd_for_2d_heisenberg_models_amd as H  # noqa: E402

HBM_PEAK_GBS = 8000.0
# the batched FP32 SMD engine's longest DSSF trajectory chunk: the
# contraction runs when a chunk fills, and the configuration-4 and -5 runs
# (8192 and 4096 points) each fit in one chunk; timed runs cover whole runs
SMD_CHUNK = 8192
# DSSF time-contraction kernel per contraction tile
GEMM_KERNEL = {0: "time_gemm_mfma_q_kernel", 1: "time_gemm_mfma_kernel",
               2: "time_gemm_mfma_q256_kernel", 3: "time_gemm_mfma_db_kernel",
               4: "time_gemm_mfma_w_kernel", 5: "time_gemm_mfma_p_kernel",
               6: "time_dft_fft_kernel"}


class HipTimer:
    """hipEvent timing on the null stream.  Library calls that work on a
    stream of their own (the replayed sweep graph) wait for it before they
    return, so the interval still covers their work."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                                 ctypes.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.a, self.b = ctypes.c_void_p(), ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(self.a)) == 0
        assert self.hip.hipEventCreate(ctypes.byref(self.b)) == 0

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0

    def start(self):
        self.sync()
        assert self.hip.hipEventRecord(self.a, None) == 0

    def stop(self):
        assert self.hip.hipEventRecord(self.b, None) == 0
        assert self.hip.hipEventSynchronize(self.b) == 0
        self.sync()
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), self.a, self.b) == 0
        return ms.value


def unit_spins(rng, shape):
    v = rng.normal(size=shape)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def mc_workload():
    c = H.Couplings(J1=1.0, J2=0.12, JX=0.05, h=0.3, lambda1=0.9, lambda2=0.9)
    lat = H.Lattice.generate("triang", 96, c)
    temps = np.geomspace(0.02, 2.0, 256)
    return c, lat, temps


def stripe_noise(rng, lat, sigma=0.15):
    """Horizontal-stripe state (spins along z, HStpPhase) with Gaussian angular
    noise: each spin is rotated about a random axis perpendicular to it."""
    L = lat.L
    n = lat.N
    s = np.zeros((n, 3))
    j = np.arange(n) % L
    s[:, 2] = np.cos(np.pi * j)
    ax = rng.normal(size=(n, 3))
    ax -= (ax * s).sum(1, keepdims=True) * s
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    ang = rng.normal(scale=sigma, size=(n, 1))
    s = s * np.cos(ang) + np.cross(ax, s) * np.sin(ang)
    return s / np.linalg.norm(s, axis=1, keepdims=True)


def smd_workload(rng):
    c = H.Couplings(J1=1.0, J2=0.6)
    lats = [H.Lattice.generate("square", 64, c), H.Lattice.generate("kagome", 32, c),
            H.Lattice.generate("hexagn", 40, c), H.Lattice.generate("square", 48, c)]
    vac = np.zeros(lats[3].N, np.int32)
    vac[rng.choice(lats[3].N, 230, replace=False)] = 1
    lats[3].set_vacancies(vac)
    idx = np.repeat(np.arange(4), 128).astype(np.int32)
    samples = []
    for i in idx:
        if i in (0, 3):
            samples.append(stripe_noise(rng, lats[i]))
        else:
            samples.append(unit_spins(rng, (lats[i].N, 3)))
    return lats, idx, samples


def mc_error_vs_fp64(lat, temps, rng):
    """Accuracy of the batched FP32 sweep against the batched FP64 engine
    on the same replicas: one overrelaxation pass of every replica (no random
    draws, so both forms update the same chain) from the same unit spins,
    and the FP64 energies each engine reports for the result.  Heat-bath
    passes draw in FP32 and FP64 differently and are compared statistically
    in the tests instead."""
    R, N = len(temps), lat.N
    s0 = unit_spins(rng, (R, N, 3))
    res = {}
    for prec in (H.BATCHED_F32, H.BATCHED_F64):
        m = H.MonteCarlo(lat, temps, precision=prec, n_or=1, n_hb=0, n_or_boost=1, n_hb_boost=0, seed=3)
        m.set_spins(s0)
        m.run(1, pt_parity=-1)
        res[prec] = (m.get_spins(), m.energies())
        del m
    (s32, e32), (s64, e64) = res[H.BATCHED_F32], res[H.BATCHED_F64]
    d = np.abs(s32 - s64).max(axis=-1)
    err = {"workload": "one overrelaxation pass of the %d replicas from the same unit spins" % R,
           "max_abs_spin_err": float(d.max()), "p9999_abs_spin_err": float(np.quantile(d, 0.9999)),
           # random spins have energies near zero, so the energy error is
           # stated per site rather than relative
           "max_energy_err_per_site": float(np.abs(e32 - e64).max() / N)}
    print("[bench] MC FP32 vs FP64 (one overrelaxation pass): max |dS| %.2e, energy %.2e per site"
          % (err["max_abs_spin_err"], err["max_energy_err_per_site"]), file=sys.stderr, flush=True)
    return err


def bench_measure(timer, mc, lat, R, sweep_ms, k):
    """Measurement-stage sweeps on the headline engine: order parameters,
    spin-stiffness bond sums and per-slot moments every sweep (obs_kernel and
    the moments in the exchange kernel), and with the static S(q) of every
    replica every 10th measured sweep (grid DFT, FP64).  Costs are differences
    against the unmeasured sweep; algorithmic bytes of obs_kernel = 12 B per
    site + 2 x 12 B per bond per replica, FLOP of one S(q) = 8 x 2 x 3 x G^3
    per replica (two separable complex DFT passes over a G x G grid)."""
    N, nb = lat.N, lat.n_bonds
    mc.run(5, pt_parity=2, measure=True)
    timer.start()
    mc.run(k, pt_parity=2, measure=True)
    m_ms = timer.stop() / k
    mc.set_sq(10)
    mc.run(10, pt_parity=2, measure=True)
    timer.start()
    mc.run(10 * (k // 10), pt_parity=2, measure=True)
    q_ms = timer.stop() / (10 * (k // 10))
    mc.set_sq(0)
    obs_ms = max(m_ms - sweep_ms, 1e-6)
    sq_ms = max(10 * (q_ms - m_ms), 1e-6)
    G = lat.Gsz
    obs_bytes = (12.0 * N + 24.0 * nb) * R
    sq_flop = 8.0 * 2 * 3 * G ** 3 * R
    print("[bench] MC measured sweep %.3f ms (observables +%.3f ms); with S(q) every 10: %.3f ms "
          "(one S(q) of %d replicas %.3f ms)" % (m_ms, obs_ms, q_ms, R, sq_ms),
          file=sys.stderr, flush=True)
    return {"ms_per_measured_sweep": m_ms, "observables_ms": obs_ms,
            "observables_gbs": obs_bytes / (obs_ms * 1e-3) / 1e9,
            "ms_per_sweep_with_sq_every_10": q_ms, "sq_ms_per_call": sq_ms,
            "sq_grid": G, "sq_tflops_fp64": sq_flop / (sq_ms * 1e-3) / 1e12,
            "sweeps": k}


def graph_workload(rng, n_target=12000):
    """Poisson-disk graph of about n_target sites (Bridson sampling fills a
    periodic box at ~0.627 points per r_min^2) with 5 % vacancies."""
    box = math.sqrt(n_target / 0.627)
    pts = H.poisson_disk(box, 1.0, seed=5)
    c = H.Couplings(J1=1.0, J2=0.4, JX=0.1)
    lat = H.Lattice.from_points(pts, box, (1.15, 1.45, 1.75), c)
    vac = np.zeros(lat.N, np.int32)
    vac[rng.choice(lat.N, int(round(0.05 * lat.N)), replace=False)] = 1
    lat.set_vacancies(vac)
    return c, lat, box


def q_path(n=200, qmax=2 * math.pi):
    """Gamma -> (qmax, 0) -> (qmax, qmax) -> Gamma, n points by arc length."""
    corners = np.array([[0.0, 0.0], [qmax, 0.0], [qmax, qmax], [0.0, 0.0]])
    seg = np.linalg.norm(np.diff(corners, axis=0), axis=1)
    s = np.linspace(0.0, seg.sum(), n)
    out = np.zeros((n, 2))
    for k, v in enumerate(s):
        i = min(np.searchsorted(np.cumsum(seg), v, side="right"), 2)
        t = (v - (np.cumsum(seg)[i - 1] if i > 0 else 0.0)) / seg[i]
        out[k] = corners[i] + min(max(t, 0.0), 1.0) * (corners[i + 1] - corners[i])
    return out


def bench_graph(timer, rng, tune, mc_steps=50):
    c, lat, box = graph_workload(rng)
    N = lat.N
    print("[bench] graph: N=%d box=%.1f Z=%d colours=%d" % (N, box, lat.zn1 + lat.zn2 + lat.znx,
                                                            lat.n_colours), file=sys.stderr, flush=True)
    temps = np.geomspace(0.05, 1.5, 128)
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=5, n_hb=1, n_or_boost=5,
                      n_hb_boost=1, pt_cap50=False, seed=5, tune=tune, tune_sweeps=5)
    mtune = mc.tuning
    mc.set_spins(unit_spins(rng, (128, N, 3)))
    mc.run(5, pt_parity=2)
    timer.start()
    mc.run(mc_steps, pt_parity=2)
    mms = timer.stop()
    mrate = 128 * N * 6 * mc_steps / (mms * 1e-3)
    del mc
    S, T = 256, 4096
    smd = H.SpinDynamics([lat], np.zeros(S, np.int32), 0.01, T, n_omega=512,
                         precision=H.BATCHED_F32, tune=tune, tune_steps=64)
    stune = smd.tuning
    init = [unit_spins(rng, (N, 3)) for _ in range(S)]
    smd.set_samples(init)
    smd.step(16)        # warm-up
    # the whole 4096-point run, its DSSF contraction included
    smd_steps = T - 1
    smd.set_samples(init)
    timer.start()
    smd.step(smd_steps)
    sms = timer.stop()
    srate = S * N * smd_steps / (sms * 1e-3)
    q = q_path(200)
    timer.start()
    d = smd.dssf_q(q)
    qms = timer.stop()
    dflop = 8.0 * 3 * len(q) * N * 512 * S
    # executed on the matrix cores: per (sample, component) block a
    # (2 x 512 bins) x (2 x 256 phase rows) x (N rounded to 32 sites) GEMM, three split products
    qflop = 2.0 * 1024 * 512 * ((N + 31) // 32 * 32) * 3 * 3 * S
    print("[bench] graph MC f32: %.3f ms/sweep %.3e spin-updates/s; SMD f32: %.3f ms/step "
          "%.3e site-steps/s; q-path DSSF %.1f ms" % (mms / mc_steps, mrate, sms / smd_steps, srate,
                                                     qms), file=sys.stderr, flush=True)
    return {"workload": "Poisson-disk graph N=%d (r_min=1, periodic box %.1f), J1=1 r<1.15, "
                        "J2=0.4 r<1.45, JX=0.1 r<1.75, 5%% vacancies" % (N, box),
            "sites": N, "colours": lat.n_colours,
            "mc": {"value": mrate, "unit": "spin-updates/s", "replicas": 128,
                   "ms_per_sweep": mms / mc_steps, "sweeps": mc_steps, "tuning": mtune},
            "smd": {"value": srate, "unit": "site-RK4-steps/s", "samples": S, "steps": smd_steps,
                    "ms_per_step": sms / smd_steps, "n_steps_total": T, "tuning": stune},
            "dssf_qpath": {"q_points": len(q), "ms": qms,
                           "tflops": dflop / (qms * 1e-3) / 1e12,
                           "executed_tflops": qflop / (qms * 1e-3) / 1e12,
                           "kernel": "time_gemm_mfma_w_kernel (sites as the reduction index)",
                           "finite": bool(np.isfinite(d).all())}}


def host_cores():
    """Host threads the CPU baselines use: the process's CPU affinity, capped
    at the 16 cores a one-GPU share of the box provides."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cpu_ref
    return cpu_ref


def cpu_baseline(lat, c, temps, seconds_target=15.0):
    cpu_ref = _oracle()
    cores = host_cores()
    R = min(len(temps), cores)
    sel = np.linspace(0, len(temps) - 1, R).round().astype(int)
    m = cpu_ref.Model(lat.tables(), (lat.zn1, lat.zn2, lat.znx), c.as_params())
    rng = np.random.default_rng(9)
    s = unit_spins(rng, (R, lat.N, 3))
    # calibrate with one sweep, then run a bounded sample
    t0 = time.perf_counter()
    m.mc_reference(0, temps[sel], s, 1, n_or=5, n_hb=1, n_or_b=5, n_hb_b=1, pt_start=0,
                   threads=R)
    t1 = time.perf_counter() - t0
    n = max(1, min(200, int(seconds_target / max(t1, 1e-6))))
    t0 = time.perf_counter()
    m.mc_reference(0, temps[sel], s, n, n_or=5, n_hb=1, n_or_b=5, n_hb_b=1, pt_start=0,
                   threads=R)
    dt = time.perf_counter() - t0
    ups = R * lat.N * 6 * n / dt
    # single core: one replica, one thread, a few seconds
    n1 = max(1, min(50, int(0.25 * seconds_target / max(t1, 1e-6))))
    t0 = time.perf_counter()
    m.mc_reference(0, temps[sel[:1]], s[:1].copy(), n1, n_or=5, n_hb=1, n_or_b=5, n_hb_b=1,
                   pt_start=0, threads=1)
    d1 = time.perf_counter() - t0
    return {"value": ups, "unit": "spin-updates/s", "cores": R, "kind": "port",
            "single_core": lat.N * 6 * n1 / d1,
            "sample": "reference-order FP64 sweeps (random permutation, 5 OR + 1 HB, bond-list "
                      "energy, replica exchange) of %d replicas of the triangular L=96 lattice "
                      "for %d sweeps, one OpenMP thread per replica; single_core: one replica, "
                      "%d sweeps, one thread" % (R, n, n1)}


def cpu_baseline_config1(seconds_target=10.0):
    """Configuration 1 on the host: square L=24, J1=1, J2=0.6, 16 replicas in
    [0.05, 1.5], 1 HB + 5 OR, reference-order FP64 chain with the exchange."""
    cpu_ref = _oracle()
    c = H.Couplings(J1=1.0, J2=0.6)
    lat = H.Lattice.generate("square", 24, c)
    temps = np.geomspace(0.05, 1.5, 16)
    m = cpu_ref.Model(lat.tables(), (4, 4, 4), c.as_params())
    s = unit_spins(np.random.default_rng(11), (16, lat.N, 3))
    cores = host_cores()
    t0 = time.perf_counter()
    m.mc_reference(0, temps, s, 20, n_or=5, n_hb=1, pt_start=0, threads=cores)
    t1 = (time.perf_counter() - t0) / 20
    n = max(20, min(10000, int(seconds_target / max(t1, 1e-9))))
    t0 = time.perf_counter()
    m.mc_reference(0, temps, s, n, n_or=5, n_hb=1, pt_start=0, threads=cores)
    dt = time.perf_counter() - t0
    return {"value": 16 * lat.N * 6 * n / dt, "unit": "spin-updates/s", "cores": min(cores, 16),
            "kind": "port", "sample": "%d sweeps of the 16 replicas, one OpenMP thread per "
                                      "replica" % n}


def cpu_baseline_smd(seconds_target=15.0):
    """The reference's SMD path on the host for configuration 2: per sample
    the oracle's reference-order RK4 with the in-place (Gauss-Seidel)
    correction every step over the square L=24 lattice, then the windowed
    time FFT of the gridded spatial DFT (numpy, the MKL calls' transform),
    one thread per sample (one MPI rank each in the reference)."""
    from concurrent.futures import ThreadPoolExecutor
    cpu_ref = _oracle()
    c = H.Couplings(J1=1.0, J2=0.6)
    lat = H.Lattice.generate("square", 24, c)
    m = cpu_ref.Model(lat.tables(), (4, 4, 4), c.as_params())
    gm = lat.grid_map()
    rng = np.random.default_rng(12)
    cores = host_cores()
    T = 4096

    def one(k, steps):
        s0 = stripe_noise(np.random.default_rng(100 + k), lat)
        _, tr, _, _ = m.smd_run(s0, 0.01, steps, gauss_seidel=True)
        cpu_ref.dssf_fft(tr, gm, 24, np.ones(steps), min(512, steps))

    t0 = time.perf_counter()
    one(0, 256)
    t1 = (time.perf_counter() - t0) / 256
    steps = max(256, min(T, int(seconds_target / max(t1, 1e-9))))
    S = cores
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        list(ex.map(lambda k: one(k, steps), range(S)))
    dt = time.perf_counter() - t0
    return {"value": S * lat.N * (steps - 1) / dt, "unit": "site-RK4-steps/s", "cores": cores,
            "kind": "port",
            "sample": "%d stripe + noise samples of configuration 2 (square L=24), %d time points "
                      "each: reference-order FP64 RK4 with the in-place correction, then the "
                      "windowed time FFT of the spatial DFT (512 bins), one thread per sample"
                      % (S, steps)}


def profile_value(name, key):
    """A `key = value` figure from a committed rocprofv3 summary in profiles/."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.split("=")[0].strip() == key:
            try:
                return float(line.split("=")[1].split()[0].replace(",", ""))
            except (ValueError, IndexError):
                return None
    return None


def read_traffic(name="mc_sweep_f32_triang96_r256_pmc.txt", key="traffic_bytes_per_launch"):
    """HBM bytes per launch (or per step) from a committed rocprofv3 --pmc summary."""
    return profile_value(name, key)


def calibrate_copy(timer, nbytes=1 << 30, reps=5):
    """Device-to-device copies of a known size: the HBM counters of this blit
    (FETCH_SIZE = WRITE_SIZE = 1 GiB per copy ideally) calibrate the counters."""
    hip = timer.hip
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(b), nbytes) == 0
    assert hip.hipMemset(a, 1, nbytes) == 0
    timer.start()
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, nbytes, 3) == 0  # hipMemcpyDeviceToDevice
    ms = timer.stop()
    print("[bench] copy %d x %d bytes: %.3f ms each, %.1f GB/s (read+write)"
          % (reps, nbytes, ms / reps, 2 * nbytes * reps / (ms * 1e-3) / 1e9), file=sys.stderr)
    hip.hipFree(a)
    hip.hipFree(b)


# MI355X: 256 CUs x 4 SIMDs; a wave64 VALU instruction occupies a 16-lane
# SIMD for 4 cycles; 2.4 GHz peak engine clock
SIMDS = 1024
VALU_PEAK_GINST = SIMDS * 2.4e9 / 4 / 1e9


def valu_roofline(profile, kernel_s, hbm_alg_bytes, kernel, unit="launch", extra=None,
                  traffic_profile=None):
    """Issue-bound roofline of an LDS-resident kernel: VALU instructions per
    unit (SQ_INSTS_VALU from the committed PMC summary) over the live time of
    one unit against the SIMDs' issue peak, with the HBM view beside it
    (algorithmic HBM bytes per unit over the same time, and the PMC HBM bytes
    per unit).  unit: "launch" (one sweep launch: kernel_s is a launch's
    duration) or "step" (one RK4 step of a persistent launch: kernel_s is a
    launch's duration divided by its steps; the profile's counts are divided
    by the steps of the profiled launch), so the two need not run equally
    many steps per launch."""
    # kernels that order their phases per wave also issue the instructions of
    # their progress-count waits: the roofline rates the arithmetic alone
    # (counted with a workgroup barrier per phase) and reports the issued
    # count beside it
    issued = profile_value(profile, "valu_insts_per_" + unit)
    vi = profile_value(profile, "valu_work_insts_per_" + unit)
    if vi is None:
        vi = issued
    traffic = profile_value(traffic_profile or profile, "traffic_bytes_per_" + unit)
    lds = profile_value(profile, "lds_wait_frac")
    per = 1.0
    ach = None if vi is None else vi / kernel_s / 1e9
    hbm_gbs = hbm_alg_bytes / kernel_s / 1e9
    out = {"bound": "valu", "achieved": ach, "peak": VALU_PEAK_GINST, "unit": "Ginst/s",
           "frac": None if ach is None else ach / VALU_PEAK_GINST,
           "traffic": traffic, "kernel": kernel, "kernel_ms": kernel_s * 1e3,
           "lds_wait_frac": lds,
           "valu_issued": None if issued is None else
           {"insts_per_" + unit: issued, "frac": issued / kernel_s / 1e9 / VALU_PEAK_GINST},
           "hbm": {"achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": hbm_gbs / HBM_PEAK_GBS,
                   "algorithmic_bytes_per_launch": hbm_alg_bytes,
                   "measured_frac": None if traffic is None else
                   traffic * per / kernel_s / 1e9 / HBM_PEAK_GBS},
           "profile": "profiles/" + profile,
           "traffic_profile": "profiles/" + (traffic_profile or profile)}
    if extra:
        out.update(extra)
    return out


def bench_mc(timer, args, rng, tune):
    """Configuration 3: the headline engine, its kernel roofline, the
    measurement stage, and the same workload in the batched FP64, plain FP64
    and Ising forms."""
    c, lat, temps = mc_workload()
    R, N = len(temps), lat.N
    ups_per_sweep = R * N * 6
    print("[bench] MC workload: triang L=96 N=%d R=%d colours=%d" % (N, R, lat.n_colours),
          file=sys.stderr, flush=True)
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=5, n_hb=1, n_or_boost=5,
                      n_hb_boost=1, pt_cap50=False, seed=3, tune=tune, tune_sweeps=5)
    mc.set_spins(unit_spins(rng, (R, N, 3)))
    mc.run(args.warmup, pt_parity=2)
    timer.start()
    mc.run(args.steps, pt_parity=2)
    ms = timer.stop()
    value = ups_per_sweep * args.steps / (ms * 1e-3)
    # dominant kernel alone: back-to-back sweep launches, no exchange kernel
    ksteps = max(10, args.steps // 2)
    timer.start()
    mc.sweeps_only(ksteps)
    kms = timer.stop() / ksteps
    z = 2 * lat.n_bonds // N          # coupled slots per site (18 here)
    # the kernel the engine launches, as the library names it (the eighteen-slot
    # float4 replica statically placed in LDS, classes ordered per wave, here)
    mc_kernel = mc.sweep_kernel()
    gather_bytes = (z + 2) * 12 + 4 * z   # own read+write, z spins, z indices
    # HBM bytes a sweep must move with each replica resident in LDS: one read
    # and one write of its spins (12 B per site) plus one pass over the packed
    # 16-bit neighbour labels ((z + 1) / 2 words per site, shared by all
    # replicas) and over the energy's half-shell labels ((z / 2 + 1) / 2 words)
    hbm_bytes = R * 2 * 12 * (N + 1) + 4 * ((z + 1) // 2 + (z // 2 + 1) // 2) * N
    att, acc = mc.pt_counters()
    roof = valu_roofline("mc_sweep_f32_triang96_r256_valu_pmc.txt", kms * 1e-3, hbm_bytes,
                         mc_kernel, traffic_profile="mc_sweep_f32_triang96_r256_pmc.txt",
                         extra={"gather_model": {
                             "bytes_per_update": gather_bytes,
                             "achieved": ups_per_sweep * gather_bytes / (kms * 1e-3) / 1e9,
                             "unit": "GB/s",
                             "note": "neighbour gathers served from LDS, not HBM"},
                             # LDS view: 20 float4 accesses per update (z neighbours,
                             # the own spin and its store)
                             "lds": {"bytes_per_launch": ups_per_sweep * 16 * (z + 2),
                                     "achieved": ups_per_sweep * 16 * (z + 2) / (kms * 1e-3) / 1e9,
                                     "unit": "GB/s"}})
    print("[bench] MC batched f32: %.3f ms/sweep, %.3e spin-updates/s, kernel %.3f ms, VALU frac %s"
          % (ms / args.steps, value, kms, roof["frac"]), file=sys.stderr, flush=True)
    out = {
        "metric": "MC spin-updates/s & SMD site-RK4-steps/s: batched mixed-prec vs plain FP64",
        "value": value,
        "unit": "spin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 spins/fields, f64 energies and exchange",
        "data": "synthetic: random unit spins, seeded",
        "config": {"workload": "batched PT precision form: triangular L=96 N=9216, J1=1 J2=0.12 "
                               "JX=0.05 h=0.3 lambda=0.9, 256 replicas T in [0.02,2] geometric, "
                               "sweep = 5 OR + 1 HB + replica exchange",
                   "replicas": R, "sites": N, "colours": lat.n_colours,
                   "pt_accept_rate": float(acc.sum() / max(1, att.sum())),
                   "tuning": mc.tuning},
        "roofline": roof,
    }
    out["measure"] = bench_measure(timer, mc, lat, R, ms / args.steps, max(10, args.steps // 4))
    out["mc_error_vs_fp64"] = mc_error_vs_fp64(lat, temps, rng)
    del mc
    # the same workload with FP64 state, all replicas in one launch
    mb = H.MonteCarlo(lat, temps, precision=H.BATCHED_F64, n_or=5, n_hb=1, n_or_boost=5,
                      n_hb_boost=1, pt_cap50=False, seed=3)
    mb.set_spins(unit_spins(rng, (R, N, 3)))
    mb.run(2, pt_parity=2)
    bs = 10
    timer.start()
    mb.run(bs, pt_parity=2)
    bms = timer.stop()
    bv = ups_per_sweep * bs / (bms * 1e-3)
    out["batched_f64"] = {"value": bv, "ms_per_step": bms / bs, "steps": bs,
                          "f32_over_f64": value / bv,
                          "form": "FP64 spins, one workgroup per replica, all replicas in one "
                                  "launch (a replica of 221 KB exceeds LDS: global-memory sweep)"}
    print("[bench] MC batched f64: %.2f ms/sweep" % (bms / bs), file=sys.stderr, flush=True)
    del mb
    if not args.no_plain:
        mp = H.MonteCarlo(lat, temps, precision=H.PLAIN_F64, n_or=5, n_hb=1, n_or_boost=5,
                          n_hb_boost=1, pt_cap50=False, seed=3)
        mp.set_spins(unit_spins(rng, (R, N, 3)))
        mp.run(1, pt_parity=2)
        ps = 3
        timer.start()
        mp.run(ps, pt_parity=2)
        pms = timer.stop()
        pv = ups_per_sweep * ps / (pms * 1e-3)
        out["plain_f64"] = {"value": pv, "ms_per_step": pms / ps, "steps": ps,
                            "batched_over_plain": value / pv,
                            "form": "FP64, one replica per launch on one workgroup (1 of 256 "
                                    "CUs), replicas launched one after another"}
        print("[bench] MC plain f64: %.1f ms/sweep" % (pms / ps), file=sys.stderr, flush=True)
        del mp
    # Ising variant: int8 spins, Metropolis, h = 0.3 on the Ising axis
    mi = H.MonteCarlo(lat, temps, model=H.ISING, precision=H.BATCHED_F32, seed=3)
    si = np.zeros((R, N, 3))
    si[:, :, 0] = rng.choice([-1.0, 1.0], size=(R, N))
    mi.set_spins(si)
    mi.run(5, pt_parity=2)
    isw = max(20, args.steps)
    timer.start()
    mi.run(isw, pt_parity=2)
    ims = timer.stop()
    # the sweep kernel alone: back-to-back launches, no exchange kernel
    iks = max(10, args.steps // 2)
    timer.start()
    mi.sweeps_only(iks)
    ikms = timer.stop() / iks
    # HBM: one read and one write of each int8 replica, the packed labels
    # ((z + 1) / 2 words per site), the counter labels and the energy's
    # half-shell labels ((z / 2 + 1) / 2 words per site)
    ihbm = R * 2 * (N + 1) + 4 * ((z + 1) // 2 + 1 + (z // 2 + 1) // 2) * N
    out["ising"] = {"value": R * N * isw / (ims * 1e-3), "unit": "spin-updates/s",
                    "ms_per_step": ims / isw, "steps": isw,
                    "workload": "configuration 3 graph, 256 replicas, one Metropolis pass + "
                                "replica exchange per sweep, int8 spins",
                    "roofline": valu_roofline("mc_ising_sweep_triang96_r256_valu_pmc.txt", ikms * 1e-3,
                                              ihbm, mi.sweep_kernel())}
    print("[bench] MC ising: %.3f ms/sweep" % (ims / isw), file=sys.stderr, flush=True)
    del mi
    return out, (c, lat, temps)


def bench_config1(timer, rng):
    """Configuration 1: square L=24 J1-J2, 16 PT temperatures in [0.05, 1.5],
    1 HB + 5 OR, 5e3 thermalisation + 5e3 measurement sweeps (the whole run)."""
    c = H.Couplings(J1=1.0, J2=0.6)
    lat = H.Lattice.generate("square", 24, c)
    temps = np.geomspace(0.05, 1.5, 16)
    units = 16 * lat.N * 6
    out = {"workload": "square L=24 J1=1 J2=0.6, 16 replicas T in [0.05,1.5] geometric, "
                       "1 HB + 5 OR, 5000 + 5000 sweeps with replica exchange"}
    for name, prec in (("plain_f64", H.PLAIN_F64), ("batched_f64", H.BATCHED_F64),
                       ("batched_f32", H.BATCHED_F32)):
        mc = H.MonteCarlo(lat, temps, precision=prec, n_or=5, n_hb=1, seed=1)
        mc.set_spins(unit_spins(rng, (16, lat.N, 3)))
        n_th, n_me = (5000, 5000) if prec != H.PLAIN_F64 else (50, 50)
        timer.start()
        mc.run(n_th, pt_parity=2)
        mc.reset_counters()
        for b in range(max(1, n_me // 1000)):
            mc.run(min(1000, n_me), pt_parity=b % 2, measure=True)
            mc.moments(reset=True)
        ms = timer.stop()
        n = n_th + max(1, n_me // 1000) * min(1000, n_me)
        out[name] = {"value": units * n / (ms * 1e-3), "ms_per_sweep": ms / n, "sweeps": n}
        del mc
    out["batched_f32_over_plain_f64"] = out["batched_f32"]["value"] / out["plain_f64"]["value"]
    print("[bench] config 1: plain f64 %.3e, batched f64 %.3e, batched f32 %.3e spin-updates/s"
          % (out["plain_f64"]["value"], out["batched_f64"]["value"], out["batched_f32"]["value"]),
          file=sys.stderr, flush=True)
    return out


def bench_smd(timer, args, rng, tune):
    lats, idx, samples = smd_workload(rng)
    sites = sum(lats[i].N for i in idx)
    k = args.smd_steps           # RK4 steps of the run: n_steps = k + 1 time points
    T = k + 1
    chunk = min(T, SMD_CHUNK)    # FP32 trajectory chunk (memory permitting)
    smd = H.SpinDynamics(lats, idx, 0.005, T, n_omega=512, precision=H.BATCHED_F32,
                         tune=tune, tune_steps=64)
    # breakdown run (also the warm-up): two steps (t = 0 recorded, kernels
    # loaded), the RK4 kernel alone over 1000 steps (one launch, no chunk
    # fills), then the step that fills the chunk (one RK4 step + the
    # contraction of the whole chunk) against a plain step
    smd.set_samples(samples)
    smd.step(2)
    nb = min(1000, chunk - 5)
    timer.start()
    smd.step(nb)
    rk_ms = timer.stop() / nb
    smd.step(chunk - 4 - nb)
    timer.start()
    smd.step(1)
    t_flush = timer.stop()
    gemm_ms = max(t_flush - rk_ms, 1e-6)
    if T > chunk:
        smd.step(T - 1 - (chunk - 1))
    # the timed run: all k steps of the configuration from the initial samples,
    # the DSSF contraction of every chunk included
    smd.set_samples(samples)
    timer.start()
    smd.step(k)
    sms = timer.stop()
    srate = sites * k / (sms * 1e-3)
    gm = sum(3 * lats[i].N for i in idx)
    mfma_flop = 2.0 * gm * 1024 * chunk * 3      # M x (2 n_omega) x K, three split products
    # algorithmic DSSF contraction FLOP per chunk: 8 x 3 x Nq x n_omega x T per sample
    alg_flop = sum(8.0 * 3 * lats[i].Nq * 512 * chunk for i in idx)
    # gather model per site-RK4-step: 6 field passes x [(z_i + 2) 12 B + 4 z_i B],
    # summed over sites: 6 (24 N + 32 n_bonds) per sample
    rk_gather = sum(6 * (24 * lats[i].N + 32 * lats[i].n_bonds) for i in idx)
    # HBM bytes one step must move with every sample resident in LDS: the
    # FP32 trajectory row the contraction reads back (3 N values per sample)
    rk_hbm = sum(4 * 3 * lats[i].N for i in idx)
    gemm = smd.tuning["gemm"] if smd.tuning else 6
    roof = valu_roofline("smd_rk4_f32_config4_pmc.txt", rk_ms * 1e-3, rk_hbm,
                         smd.step_kernel() + " (%d steps per launch)" % nb,
                         unit="step",
                         extra={"per": "RK4 step of the 512-sample batch",
                                "gather_model": {"bytes_per_step": rk_gather,
                                                 "achieved": rk_gather / (rk_ms * 1e-3) / 1e9,
                                                 "unit": "GB/s",
                                                 "note": "neighbour gathers served from LDS"}})
    # the chunk's time transform: the trajectory chunk read once, Y (3 N rows x
    # 2 n_omega FP32 per sample) written once
    tr_bytes = 4 * gm * chunk + 4 * gm * 1024
    if gemm == 6:
        # the pruned FFT: VALU-issue bound (the profile's VALU instructions
        # per launch over the live launch time), HBM beside it
        troof = valu_roofline("dssf_fft_config4_pmc.txt", gemm_ms * 1e-3, tr_bytes,
                              GEMM_KERNEL[6] + " (one %d-point chunk)" % chunk,
                              extra={"note": "T / 512 decimated 512-point FFTs per row pair "
                                             "and one complex multiply-add per bin and t2; "
                                             "HBM: the chunk read once, Y written once"})
    else:
        troof = {"bound": "mfma",
                 "achieved": alg_flop / (gemm_ms * 1e-3) / 1e12,
                 "peak": 2500.0, "unit": "TFLOP/s",
                 "frac": alg_flop / (gemm_ms * 1e-3) / 2.5e15,
                 "traffic": read_traffic("dssf_contraction_mfma_pmc.txt"),
                 "kernel": GEMM_KERNEL.get(gemm, "?"),
                 "algorithmic_flop_per_launch": alg_flop,
                 "executed_flop_per_launch": mfma_flop,
                 "executed_tflops": mfma_flop / (gemm_ms * 1e-3) / 1e12,
                 "mfma_busy_frac": profile_value("dssf_contraction_mfma_pmc.txt", "mfma_busy_frac"),
                 "profile": "profiles/dssf_contraction_mfma_pmc.txt"}
    out = {"value": srate, "unit": "site-RK4-steps/s", "steps": k,
           "ms_per_step": sms / k, "samples": len(idx), "sites": sites, "n_steps_total": T,
           "workload": "512 samples: square L=64, kagome L=32, hexagonal L=40, "
                       "square L=48 + 10% vacancies; FP32 RK4 dt=0.005, FP64 "
                       "correction, DSSF time transform over 512 bins included",
           "tuning": smd.tuning, "roofline": roof,
           "dssf_chunk_transform": {"ms": gemm_ms, "time_points": chunk, "form": gemm,
                                    "roofline": troof}}
    print("[bench] SMD batched f32: %.4f ms/step, %.3e site-steps/s, RK4 kernel %.4f ms/step, "
          "chunk transform %.2f ms (%s)"
          % (sms / k, srate, rk_ms, gemm_ms, GEMM_KERNEL.get(gemm, "?")),
          file=sys.stderr, flush=True)
    del smd
    # the same batch with FP64 state (global-memory form, FP64 GEMM contraction)
    sb = H.SpinDynamics(lats, idx, 0.005, 520, n_omega=512, precision=H.BATCHED_F64)
    sb.set_samples(samples)
    sb.step(3)
    fs = 256
    timer.start()
    sb.step(fs)    # 256 steps with one 256-point FP64 chunk contraction among them
    fms = timer.stop()
    out["batched_f64"] = {"value": sites * fs / (fms * 1e-3), "ms_per_step": fms / fs,
                          "steps": fs, "f32_over_f64": srate / (sites * fs / (fms * 1e-3)),
                          "form": "FP64 state and FP64 GEMM contraction, all samples in one "
                                  "launch (global-memory form)"}
    print("[bench] SMD batched f64: %.4f ms/step" % (fms / fs), file=sys.stderr, flush=True)
    del sb
    out["dssf_error_vs_fp64"] = dssf_error(rng, lats, samples, idx)
    return out


def dssf_error(rng, lats, samples, idx):
    """Accuracy of the FP32 form's DSSF against the FP64 form on a subset of
    the configuration-4 batch (two samples each of square L=48 with its
    vacancies and kagome L=32), 2560 time points (one chunk), dt = 0.002,
    512 bins: the integrated runs, FP32 RK4 with the pruned-FFT time
    transform (the configuration-4 default) and with the split-FP16 MFMA
    contraction (128 x 512 tile), each against FP64 RK4 + FP64 GEMM."""
    sub = [lats[3], lats[1]]
    pick = [int(np.nonzero(idx == 3)[0][0]), int(np.nonzero(idx == 1)[0][0]),
            int(np.nonzero(idx == 3)[0][1]), int(np.nonzero(idx == 1)[0][1])]
    sidx = np.array([0, 1, 0, 1], np.int32)
    ss = [samples[k] for k in pick]
    T, no, dt = 2560, 512, 0.002
    res = {}
    for prec, gemm in ((H.BATCHED_F64, None), (H.BATCHED_F32, 6), (H.BATCHED_F32, 4)):
        e = H.SpinDynamics(sub, sidx, dt, T, n_omega=no, precision=prec)
        if gemm is not None:
            e.set_shape(0, gemm)
        e.set_samples(ss)
        e.step(T - 1)
        res[gemm] = [e.dssf(li)[0] for li in range(2)]
    out = {"subset": "2 x square L=48 (230 vacancies) + 2 x kagome L=32, 2560 points, dt=0.002, "
                     "512 bins", "floor": "bins above 1e-2 of their (q, component) peak"}
    for gemm, key in ((6, "integrated"), (4, "integrated_mfma_tile")):
        pk, rl = [], []
        for li in range(2):
            a, b = res[None][li], res[gemm][li]
            pk.append(np.abs(a - b).max() / a.max())
            big = a > 1e-2 * a.max(axis=2, keepdims=True)
            rl.append((np.abs(a - b)[big] / a[big]).max())
        out[key] = {"form": GEMM_KERNEL[gemm], "max_err_over_peak": float(max(pk)),
                    "max_rel_err_above_floor": float(max(rl))}
        print("[bench] DSSF FP32 (%s) vs FP64: peak-normalised %.2e, relative above floor %.2e"
              % (GEMM_KERNEL[gemm], max(pk), max(rl)), file=sys.stderr, flush=True)
    return out


def bench_config2(timer, rng):
    c2 = H.Couplings(J1=1.0, J2=0.6)
    l2 = H.Lattice.generate("square", 24, c2)
    s2 = [stripe_noise(rng, l2) for _ in range(32)]
    out = {"workload": "32 stripe + noise samples, square L=24, dt=0.01, 4096 points, 512 bins"}
    for name, prec, k in (("plain_f64", H.PLAIN_F64, 64), ("batched_f64", H.BATCHED_F64, 256),
                          ("batched_f32", H.BATCHED_F32, 256)):
        sp = H.SpinDynamics([l2], np.zeros(32, np.int32), 0.01, 4096, n_omega=512, precision=prec)
        sp.set_samples(s2)
        sp.step(4)
        timer.start()
        sp.step(k)
        ms = timer.stop()
        out[name] = {"value": 32 * l2.N * k / (ms * 1e-3), "ms_per_step": ms / k, "steps": k}
        del sp
    out["batched_f32_over_plain_f64"] = out["batched_f32"]["value"] / out["plain_f64"]["value"]
    print("[bench] SMD config 2: plain f64 %.3e, batched f64 %.3e, batched f32 %.3e"
          % (out["plain_f64"]["value"], out["batched_f64"]["value"], out["batched_f32"]["value"]),
          file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--smd-steps", type=int, default=SMD_CHUNK - 1,
                    help="RK4 steps of the configuration-4 run (8191: its 8192 time points)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-smd", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the configuration-1 and configuration-2 lines")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuners (default launch shapes)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 5 device-to-device copies of 1 GiB (PMC byte calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engines run on one GPU (replica "
              "blocks and sample sets would shard across GPUs, but no multi-GPU path "
              "is built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.smd_steps < 1000:
        print("--smd-steps must be >= 1000", file=sys.stderr)
        sys.exit(2)
    if H.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    timer = HipTimer()
    if args.calibrate_copy:
        calibrate_copy(timer)
        return
    rng = np.random.default_rng(1)
    tune = not args.no_tune
    out, (c, lat, temps) = bench_mc(timer, args, rng, tune)
    if not args.no_extra:
        out["config1"] = bench_config1(timer, rng)
    if not args.no_smd:
        out["smd"] = bench_smd(timer, args, rng, tune)
        if not args.no_extra:
            out["smd"]["config2"] = bench_config2(timer, rng)
    if not args.no_graph:
        out["graph"] = bench_graph(timer, rng, tune)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(lat, c, temps)
        if not args.no_extra:
            out["cpu_baseline"]["config1"] = cpu_baseline_config1()
        if not args.no_smd:
            out["smd"]["cpu_baseline"] = cpu_baseline_smd()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
