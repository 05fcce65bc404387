"""Benchmark of the MC sweep engine and the SMD engine on one MI355X.

Headline workload (value): batched parallel tempering, precision form --
triangular L=96 (N=9216), J1=1, J2=0.12, JX=0.05, h=0.3, XXZ lambda=0.9,
256 replicas geometric in T in [0.02, 2.0], FP32 spins and fields, FP64
energies and exchange acceptance, one sweep = 5 overrelaxation + 1 heat-bath
passes of every replica followed by the on-device replica exchange.  A step
is one such sweep of all replicas; value = spin-updates/s.

Also reported on the same line:
  plain_f64      -- the same workload in the plain FP64 one-instance-per-launch
                    form (a few sweeps), and the batched/plain ratio;
  smd            -- batched uneven SMD: 512 samples over square L=64, kagome
                    L=32, hexagonal L=40 and square L=48 with 10 % vacancies,
                    FP32 RK4 (dt=0.005) with FP64 energy correction and the
                    DSSF time contraction (512 frequency bins) included, in
                    site-RK4-steps/s; and the plain FP64 SMD form on the
                    L=24 square lattice with 32 samples;
  roofline       -- the sweep kernel's algorithmic bytes per launch over its
                    average launch duration (HIP events on the launch stream);
  graph          -- irregular graph: about 12,000 seeded Poisson-disk points
                    (r_min = 1) in a periodic box, shells J1=1 (r<1.15),
                    J2=0.4 (<1.45), JX=0.1 (<1.75), 5 % vacancies; MC with
                    128 FP32 replicas, SMD with 256 samples over 4096 RK4
                    steps (dt=0.01) and the DSSF along a 200-point q-path;
  cpu_baseline   -- the oracle's reference-order FP64 sweeps (one thread per
                    replica, like one MPI rank each) on a bounded sample.
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
# time points per DSSF trajectory chunk of the batched FP32 SMD engine: the
# contraction runs when a chunk fills, so timed windows hold whole chunks
SMD_CHUNK = 2048
# RK4 kernel of the configuration-4 batch (slot classes 8 and 12) per SMD form
SMD_FORM_KERNEL = {1: "smd_kernel<float,4,true,12,false>", 2: "smd_kernel_two<float,4,8,12>",
                   3: "smd_kernel_wide<float,12,false>"}


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


def bench_graph(timer, rng, tune, mc_steps=50, smd_steps=SMD_CHUNK):
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
    smd.set_samples([unit_spins(rng, (N, 3)) for _ in range(S)])
    smd.step(SMD_CHUNK - 1)
    timer.start()
    smd.step(smd_steps)
    sms = timer.stop()
    srate = S * N * smd_steps / (sms * 1e-3)
    rest = T - 1 - (SMD_CHUNK - 1) - smd_steps
    if rest > 0:
        smd.step(rest)
    q = q_path(200)
    timer.start()
    d = smd.dssf_q(q)
    qms = timer.stop()
    dflop = 8.0 * 3 * len(q) * N * 512 * S
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
                           "finite": bool(np.isfinite(d).all())}}


def cpu_baseline(lat, c, temps, seconds_target=15.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cpu_ref
    cores = os.cpu_count() or 1
    R = min(len(temps), max(1, cores))
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


def read_traffic(name="mc_sweep_f32_triang96_r256_pmc.txt", key="traffic_bytes_per_launch"):
    """HBM bytes per launch (or per step) from a committed rocprofv3 --pmc summary."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith(key):
            try:
                return float(line.split("=")[1])
            except ValueError:
                return None
    return None


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--smd-steps", type=int, default=SMD_CHUNK)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-smd", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
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
    if H.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    timer = HipTimer()
    if args.calibrate_copy:
        calibrate_copy(timer)
        return
    rng = np.random.default_rng(1)
    c, lat, temps = mc_workload()
    R, N = len(temps), lat.N
    ups_per_sweep = R * N * 6
    print("[bench] MC workload: triang L=96 N=%d R=%d colours=%d" % (N, R, lat.n_colours),
          file=sys.stderr, flush=True)
    tune = not args.no_tune
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
    thr = mc.tuning["threads"] if mc.tuning else 1024
    mc_kernel = ("heis_sweep_fast<float,20>" if np.bincount(lat.colours()).max() <= thr
                 else "heis_sweep_kernel<float,true,20>") + " x %d threads" % thr
    z = 2 * lat.n_bonds // N          # coupled slots per site (18 here)
    bytes_per_update = (z + 2) * 12 + 4 * z   # own read+write, z spins, z indices
    achieved = ups_per_sweep * bytes_per_update / (kms * 1e-3) / 1e9
    traffic = read_traffic()
    att, acc = mc.pt_counters()
    print("[bench] MC batched f32: %.3f ms/sweep, %.3e spin-updates/s, kernel %.3f ms"
          % (ms / args.steps, value, kms), file=sys.stderr, flush=True)
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
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": mc_kernel,
                     "bytes_per_update": bytes_per_update, "kernel_ms": kms},
    }
    out["measure"] = bench_measure(timer, mc, lat, R, ms / args.steps, max(10, args.steps // 4))
    del mc
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
                            "batched_over_plain": value / pv}
        print("[bench] MC plain f64: %.1f ms/sweep" % (pms / ps), file=sys.stderr, flush=True)
        del mp
    if not args.no_smd:
        lats, idx, samples = smd_workload(rng)
        sites = sum(lats[i].N for i in idx)
        T = 8192
        smd = H.SpinDynamics(lats, idx, 0.005, T, n_omega=512, precision=H.BATCHED_F32,
                             tune=tune, tune_steps=64)
        smd.set_samples(samples)
        smd.step(SMD_CHUNK - 1)  # warm-up: time points 0..2047, one full trajectory chunk
        k = args.smd_steps
        timer.start()
        smd.step(k)     # k RK4 steps; every SMD_CHUNK points the DSSF chunk GEMM runs
        sms = timer.stop()
        srate = sites * k / (sms * 1e-3)
        # complete the current trajectory chunk (its flush runs the GEMM)
        fill = k % SMD_CHUNK
        if fill:
            smd.step(SMD_CHUNK - fill)
        # the RK4 kernel alone: one launch advances every sample 2047 steps
        # inside the next chunk (no flush)
        timer.start()
        smd.step(SMD_CHUNK - 1)
        rk_ms = timer.stop() / (SMD_CHUNK - 1)
        # the chunk GEMM alone: the step that fills the chunk (one RK4 step +
        # the flush) against a plain step
        timer.start()
        smd.step(1)
        t_flush = timer.stop()
        timer.start()
        smd.step(1)
        t_one = timer.stop()
        gemm_ms = max(t_flush - t_one, 1e-6)
        gm = sum(3 * lats[i].N for i in idx)
        mfma_flop = 2.0 * gm * 1024 * SMD_CHUNK * 3      # M x (2 n_omega) x K, three split products
        # algorithmic DSSF contraction FLOP per chunk: 8 x 3 x Nq x n_omega x T per sample
        alg_flop = sum(8.0 * 3 * lats[i].Nq * 512 * SMD_CHUNK for i in idx)
        # gather model per site-RK4-step: 6 field passes x [(z_i + 2) 12 B + 4 z_i B],
        # summed over sites: 6 (24 N + 32 n_bonds) per sample
        rk_bytes = sum(6 * (24 * lats[i].N + 32 * lats[i].n_bonds) for i in idx)
        rk_gbs = rk_bytes / (rk_ms * 1e-3) / 1e9
        out["smd"] = {"value": srate, "unit": "site-RK4-steps/s", "steps": k,
                      "ms_per_step": sms / k, "samples": len(idx), "sites": sites,
                      "workload": "512 samples: square L=64, kagome L=32, hexagonal L=40, "
                                  "square L=48 + 10% vacancies; FP32 RK4 dt=0.005, FP64 "
                                  "correction, DSSF time contraction over 512 bins included",
                      "tuning": smd.tuning,
                      "roofline": {"bound": "hbm", "achieved": rk_gbs, "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": rk_gbs / HBM_PEAK_GBS,
                                   "traffic": read_traffic("smd_rk4_f32_config4_pmc.txt",
                                                           "traffic_bytes_per_step"),
                                   "traffic_unit": "HBM bytes per RK4 step of the batch",
                                   "kernel": SMD_FORM_KERNEL.get(
                                       smd.tuning["form"] if smd.tuning else 2, "?")
                                   + " (%d steps per launch)" % (SMD_CHUNK - 1),
                                   "bytes_per_step": rk_bytes, "kernel_ms_per_step": rk_ms},
                      "dssf_chunk_gemm": {"ms": gemm_ms, "time_points": SMD_CHUNK,
                                          "roofline": {"bound": "mfma",
                                                       "achieved": alg_flop / (gemm_ms * 1e-3) / 1e12,
                                                       "peak": 2500.0, "unit": "TFLOP/s",
                                                       "frac": alg_flop / (gemm_ms * 1e-3) / 2.5e15,
                                                       "traffic": read_traffic(
                                                           "dssf_contraction_mfma_pmc.txt"),
                                                       "kernel": "time_gemm_mfma_q_kernel",
                                                       "algorithmic_flop_per_launch": alg_flop,
                                                       "executed_flop_per_launch": mfma_flop,
                                                       "executed_tflops":
                                                           mfma_flop / (gemm_ms * 1e-3) / 1e12}}}
        print("[bench] SMD batched f32: %.3f ms/step, %.3e site-steps/s, RK4 kernel %.4f ms/step "
              "(%.0f GB/s gather model), chunk GEMM %.2f ms (%.0f TFLOP/s)"
              % (sms / k, srate, rk_ms, rk_gbs, gemm_ms, mfma_flop / (gemm_ms * 1e-3) / 1e12),
              file=sys.stderr, flush=True)
        del smd
        c2 = H.Couplings(J1=1.0, J2=0.6)
        l2 = H.Lattice.generate("square", 24, c2)
        s2 = [stripe_noise(rng, l2) for _ in range(32)]
        sp = H.SpinDynamics([l2], np.zeros(32, np.int32), 0.01, 4096, n_omega=512,
                            precision=H.PLAIN_F64)
        sp.set_samples(s2)
        sp.step(4)
        k2 = 32
        timer.start()
        sp.step(k2)
        pms = timer.stop()
        prate = 32 * l2.N * k2 / (pms * 1e-3)
        sb = H.SpinDynamics([l2], np.zeros(32, np.int32), 0.01, 4096, n_omega=512,
                            precision=H.BATCHED_F32)
        sb.set_samples(s2)
        sb.step(4)
        timer.start()
        sb.step(256)
        bms = timer.stop()
        brate = 32 * l2.N * 256 / (bms * 1e-3)
        out["smd"]["plain_f64_square24_s32"] = {"value": prate, "ms_per_step": pms / k2}
        out["smd"]["batched_f32_square24_s32"] = {"value": brate, "ms_per_step": bms / 256,
                                                  "batched_over_plain": brate / prate}
        print("[bench] SMD square24: plain f64 %.3e, batched f32 %.3e" % (prate, brate),
              file=sys.stderr, flush=True)
    if not args.no_graph:
        out["graph"] = bench_graph(timer, rng, tune)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(lat, c, temps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
