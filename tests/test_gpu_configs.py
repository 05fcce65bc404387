"""GPU parity tests at the sizes of the five benchmark workloads, through the
C ABI, against the oracle:

  1  square L=24 J1-J2 (J2=0.6), 16 PT temperatures in [0.05, 1.5],
     1 HB + 5 OR, FP64 and FP32 -- replay with exchange and the (2, 8) pass
     boost, and the observables against the reference-order chain;
  2  32 stripe + noise samples on that lattice, FP64 SMD, dt=0.01,
     4096 time points, 512-bin DSSF;
  3  triangular L=96 (J1=1, J2=0.12, JX=0.05, h=0.3, lambda=0.9), 256 FP32
     replicas in [0.02, 2.0], and its Ising variant bit-exact;
  4  square L=64, kagome L=32, hexagonal L=40 and square L=48 with 230
     vacancies in one FP32 SMD batch (four sites per thread, per-workgroup
     slot classes), and the MFMA DSSF contraction across trajectory chunks;
  5  the ~12,000-site Poisson-disk graph with 5 % vacancies: MC, the
     global-memory SMD form and the 200-point q-path DSSF.

The reference loops these cover: CF/mc_hm_code.cpp:2778-3130 (sweeps and
exchange), CF/subrouts.cpp:11734-11927 (RK4, correction, qSS) and
12062-12105 (time transform)."""
import math
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import classical_mc_and_smd_for_2d_heisenberg_models_amd as H
import cpu_ref

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-cfg]", *a, flush=True)
    sys.stdout.flush()


def _unit(rng, shape):
    v = rng.normal(size=shape)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _stripe_noise(rng, lat, sigma=0.15):
    """Horizontal stripe along z (HStpPhase, CF/subrouts.cpp:4371) with
    Gaussian angular noise: each spin rotated about a random axis
    perpendicular to it by an angle ~ N(0, sigma)."""
    n, L = lat.N, lat.L
    s = np.zeros((n, 3))
    s[:, 2] = np.cos(np.pi * (np.arange(n) % L))
    ax = rng.normal(size=(n, 3))
    ax -= (ax * s).sum(1, keepdims=True) * s
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    ang = rng.normal(scale=sigma, size=(n, 1))
    s = s * np.cos(ang) + np.cross(ax, s) * np.sin(ang)
    return s / np.linalg.norm(s, axis=1, keepdims=True)


def _holes(lat, vac):
    """The lattice's table with every entry of or towards a vacant site a hole."""
    nb = lat.tables().copy()
    v = np.nonzero(vac)[0]
    nb[v, :] = -1
    nb[np.isin(nb, v)] = -1
    return nb


def _model(lat, c, vac=None):
    nb = lat.tables() if vac is None else _holes(lat, vac)
    return cpu_ref.Model(nb, (lat.zn1, lat.zn2, lat.znx), c.as_params())


def _pmap(fn, items, workers=8):
    # the oracle's ctypes calls release the GIL
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(fn, items))


def _stepwise_replay(mc, m, lat, temps, s, n_sweeps, tol, n_or, n_hb, n_or_b, n_hb_b,
                     cap50=True, seed=1):
    """n_sweeps sweeps with the exchange after each (alternating parity), one
    at a time: before each sweep the oracle takes the engine's configurations
    and exchange state and replays that one sweep; spins (within tol), the
    energy trace, the slot map and the exchange counters must match.  One
    sweep per comparison keeps it within the horizon over which rounding
    differences (fused multiply-adds on the device, separate roundings in the
    oracle; the device's and the host's log, exp and sincos) stay near
    rounding level: sequential overrelaxation is a chaotic map and amplifies
    them exponentially over many passes.  Within one sweep a short tail
    remains (measured on configuration 1: a cascade through one
    high-temperature replica reaching 3.4e-7, the 99 % quantile of all sites
    at 1.5e-9).  tol bounds the 99 % quantile of the per-site differences,
    1e-5 their maximum.  Returns the
    largest spin difference and the number of sweeps that ran with boosted
    slots."""
    R = len(temps)
    st = (np.arange(R, dtype=np.int32), np.zeros(R, np.int64), np.zeros(R, np.int64),
          np.zeros(R, np.int32))
    cfg = s.copy()
    col = lat.colours()
    worst, boosted = 0.0, 0
    for k in range(n_sweeps):
        boosted += int(st[3].sum() > 0)
        tr = mc.run(1, pt_parity=2, trace=True)
        ref = cfg.copy()
        rtr, st = m.mc_replay(0, temps, col, ref, 1, n_or=n_or, n_hb=n_hb, n_or_b=n_or_b,
                              n_hb_b=n_hb_b, cap50=cap50, seed=seed, sweep0=k, pt_parity=2,
                              state=st)
        att, acc = mc.pt_counters()
        np.testing.assert_array_equal(att, st[1])
        np.testing.assert_array_equal(acc, st[2])
        np.testing.assert_array_equal(mc.slot_map(), st[0])
        got = mc.get_spins()
        d = np.abs(got - ref[st[0]]).max(axis=-1)
        err, q = d.max(), np.quantile(d, 0.99)
        worst = max(worst, err)
        if err > 1e-8:
            r, i = np.unravel_index(np.argmax(d), d.shape)
            _say("sweep %d: max |dS| %.2e at slot %d site %d (boosted %d), 99 %% quantile %.2e"
                 % (k, err, r, i, int(st[3][r]), q))
        # the tail sits on sites where the update is ill-conditioned (a field
        # that nearly cancels, a steep part of the heat-bath inverse CDF)
        assert q < tol and err < 1e-5, (k, err, q)
        # energies follow the spins: a replica carrying the tail moves by ~1e-9
        np.testing.assert_allclose(tr, rtr, rtol=1e-8)
        cfg = np.empty_like(got)
        cfg[st[0]] = got
    return worst, boosted


# ------------------------------------------------------------ config 1

def _config1():
    c = H.Couplings(J1=1.0, J2=0.6)
    return c, H.Lattice.generate("square", 24, c), np.geomspace(0.05, 1.5, 16)


@pytest.mark.parametrize("prec,cap", [(H.PLAIN_F64, True), (H.BATCHED_F64, True),
                                      (H.BATCHED_F64, False)])
def test_config1_fp64_replay_with_exchange_and_pass_boost(prec, cap):
    """Twenty FP64 sweeps of the 16 replicas (5 OR + 1 HB) with the exchange
    after every sweep (alternating parity, with and without the 50 % cap):
    each sweep against the oracle's colour replay, including the sweeps run
    with the (2 OR, 8 HB) composition after accepted swaps
    (CF/mc_hm_code.cpp:2781-2794)."""
    _say("config1 fp64 replay", prec, cap)
    c, lat, temps = _config1()
    m = _model(lat, c)
    rng = np.random.default_rng(101)
    s = _unit(rng, (16, lat.N, 3))
    mc = H.MonteCarlo(lat, temps, precision=prec, n_or=5, n_hb=1, n_or_boost=2, n_hb_boost=8,
                      pt_cap50=cap, seed=31)
    mc.set_spins(s)
    worst, boosted = _stepwise_replay(mc, m, lat, temps, s, 20, 1e-8, 5, 1, 2, 8, cap50=cap,
                                      seed=31)
    att, acc = mc.pt_counters()
    _say("accepted swaps %d, boosted sweeps %d, max |dS| %.2e" % (acc.sum() // 2, boosted, worst))
    assert boosted >= 1


def _binned(x, nb):
    b = np.asarray(x, dtype=np.float64).reshape(nb, -1).mean(axis=1)
    return b.mean(), b.std(ddof=1) / math.sqrt(nb)


def test_config1_observables_within_error_bars_of_reference_chain():
    """Configuration 1 with exchange on: E, C, |M|, the stripe order
    parameter and the stripe S(q) peak per temperature slot of the GPU FP32
    colour-order chain against the oracle's reference-order FP64 chain
    (random-permutation sweeps, same composition, exchange with the cap and
    the boost), 20 bins of 500 sweeps each after 2000.  Every observable
    agrees within 4 sigma and at least 85 % within 2 sigma."""
    _say("config1 observables vs reference chain")
    c, lat, temps = _config1()
    N, R = lat.N, len(temps)
    L = lat.L
    m = _model(lat, c)
    nbins, per_bin, every, therm = 20, 500, 10, 2000
    rng = np.random.default_rng(202)
    s0 = _unit(rng, (R, N, 3))
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=5, n_hb=1, n_or_boost=2,
                      n_hb_boost=8, pt_cap50=True, seed=41)
    mc.set_spins(s0)
    mc.set_sq(every)
    mc.run(therm, pt_parity=2)
    gpu = {k: [] for k in ("E", "E2", "M", "S", "sq")}
    for b in range(nbins):
        mc.moments(reset=True)
        mc.run(per_bin, pt_parity=2, measure=True)
        mom, n = mc.moments()
        gpu["E"].append(mom[:, 0] / n / N)
        gpu["E2"].append(mom[:, 1] / n / N ** 2)
        gpu["M"].append(mom[:, 2] / n)
        gpu["S"].append(mom[:, 10] / n)
        gpu["sq"].append((mom[:, 26] + mom[:, 27]) / mom[:, 65])
    gpu = {k: np.array(v) for k, v in gpu.items()}
    att, acc = mc.pt_counters()
    _say("gpu exchange rate", acc.sum() / att.sum())
    # oracle chain, measured every `every` sweeps, exchange state carried over
    ref = s0.copy()
    st = (np.arange(R, dtype=np.int32), np.zeros(R, np.int64), np.zeros(R, np.int64),
          np.zeros(R, np.int32))
    m.mc_reference(0, temps, ref, therm, n_or=5, n_hb=1, n_or_b=2, n_hb_b=8, pt_start=0,
                   seed=77, state=st)
    nb, rv, gm, zn = cpu_ref.lattice("square", L)
    cpu = {k: [] for k in ("E", "E2", "M", "S", "sq")}
    sweep = therm
    ok = gm >= 0
    for t in range(nbins * per_bin // every):
        m.mc_reference(0, temps, ref, every, n_or=5, n_hb=1, n_or_b=2, n_hb_b=8, pt_start=0,
                       seed=77, state=st, sweep0=sweep)
        sweep += every
        cos = st[0]
        sl = ref[cos]                      # slot order
        E = m.energies(sl) / N
        grid = np.zeros((R, 3, L * L))
        grid[:, :, ok] = sl[:, gm[ok], :].transpose(0, 2, 1)
        sq = np.fft.fft2(grid.reshape(R, 3, L, L)).reshape(R, 3, L * L) / L
        row = {"E": E, "E2": E ** 2, "M": [], "S": [], "sq": []}
        for k in range(R):
            o = cpu_ref.observables("square", L, nb, zn, 0, sl[k])
            pk, _ = cpu_ref.corr_len(sq[k], L, False, 2 * math.pi / L)
            row["M"].append(o[0])
            row["S"].append(o[6])
            row["sq"].append(pk[0, 1] + pk[0, 2])
        for kk in cpu:
            cpu[kk].append(row[kk])
    cpu = {k: np.array(v) for k, v in cpu.items()}
    _say("oracle exchange rate", st[2].sum() / st[1].sum())
    within2 = total = 0
    worst = 0.0
    for k in range(R):
        beta = 1.0 / temps[k]
        for name in ("E", "M", "S", "sq", "C"):
            if name == "C":
                gc = N * beta ** 2 * (gpu["E2"][:, k] - gpu["E"][:, k] ** 2)
                cc = N * beta ** 2 * (cpu["E2"][:, k].reshape(nbins, -1).mean(1)
                                      - cpu["E"][:, k].reshape(nbins, -1).mean(1) ** 2)
                ga, ge = gc.mean(), gc.std(ddof=1) / math.sqrt(nbins)
                ca, ce = cc.mean(), cc.std(ddof=1) / math.sqrt(nbins)
            else:
                ga, ge = _binned(gpu[name][:, k], nbins)
                ca, ce = _binned(cpu[name][:, k], nbins)
            sig = math.hypot(ge, ce) + 1e-12
            z = abs(ga - ca) / sig
            worst = max(worst, z)
            _say("T=%.3f %-2s gpu %.6f +- %.6f  ref %.6f +- %.6f  (%.2f sigma)"
                 % (temps[k], name, ga, ge, ca, ce, z))
            assert z < 4.0, (temps[k], name, ga, ca, sig)
            within2 += z < 2.0
            total += 1
    _say("within 2 sigma: %d / %d, worst %.2f sigma" % (within2, total, worst))
    assert within2 >= 0.85 * total


# ------------------------------------------------------------ config 2

def _config2(S=32):
    c = H.Couplings(J1=1.0, J2=0.6)
    lat = H.Lattice.generate("square", 24, c)
    rng = np.random.default_rng(2)
    return c, lat, [_stripe_noise(rng, lat) for _ in range(S)]


@pytest.mark.parametrize("prec", [H.PLAIN_F64, H.BATCHED_F64])
def test_config2_fp64_trajectories_over_100_steps(prec):
    """All 32 samples, 100 RK4 steps with the correction every step: within
    1e-10 of the oracle in the same (Jacobi) correction order, and within the
    1e-6 bar of the reference's in-place (Gauss-Seidel) order."""
    _say("config2 trajectories", prec)
    c, lat, samples = _config2()
    m = _model(lat, c)
    T = 101
    smd = H.SpinDynamics([lat], np.zeros(32, np.int32), 0.01, T, n_omega=32, precision=prec,
                         dssf=False)
    smd.set_samples(samples)
    smd.step(T - 1)
    out = smd.get_samples()
    E = smd.energies()
    worst_j = worst_gs = 0.0
    for k, s0 in enumerate(samples):
        sj, _, _, ej = m.smd_run(s0, 0.01, T, gauss_seidel=False, traj=False)
        sg = m.smd_run(s0, 0.01, T, gauss_seidel=True, traj=False)[0]
        worst_j = max(worst_j, np.abs(out[k] - sj).max())
        worst_gs = max(worst_gs, np.abs(out[k] - sg).max())
        assert E[k] == pytest.approx(ej[-1], rel=1e-10)
    _say("max |dS| jacobi %.2e, reference order %.2e" % (worst_j, worst_gs))
    assert worst_j < 1e-10
    assert worst_gs < 1e-6


@pytest.mark.parametrize("prec,win", [(H.PLAIN_F64, "nowin"), (H.BATCHED_F64, "lancz")])
def test_config2_dssf_576_q_512_bins_matches_oracle(prec, win):
    """The full configuration-2 run: 32 samples, 4096 time points, the
    DSSF at all 576 q-points and 512 frequency bins (sample-summed) and qSS,
    against the oracle's trajectories and transform (FFT form, pinned to
    orc_dssf).  Tolerance 1e-7 of the peak: two FP64 trajectories of 41 time
    units stay that close; measured value printed."""
    _say("config2 dssf", prec, win)
    c, lat, samples = _config2()
    m = _model(lat, c)
    T, no, dt = 4096, 512, 0.01
    w = cpu_ref.window(win, T)
    smd = H.SpinDynamics([lat], np.zeros(32, np.int32), dt, T, n_omega=no, precision=prec,
                         window=w)
    smd.set_samples(samples)
    smd.step(T - 1)
    d, q = smd.dssf(0)
    gm = lat.grid_map()

    def one(s0):
        _, tr, qq, _ = m.smd_run(s0, dt, T, gauss_seidel=False)
        return cpu_ref.dssf_fft(tr, gm, lat.Gsz, w, no), qq

    res = _pmap(one, samples)
    ref = sum(r[0] for r in res)
    qref = sum(r[1] for r in res)
    err = np.abs(d - ref).max() / ref.max()
    _say("dssf max err / peak %.2e, qss max err %.2e" % (err, np.abs(q - qref).max()))
    assert err < 1e-7
    np.testing.assert_allclose(q, qref, rtol=0, atol=1e-9 * 32)


# ------------------------------------------------------------ config 3

def _config3():
    c = H.Couplings(J1=1.0, J2=0.12, JX=0.05, h=0.3, lambda1=0.9, lambda2=0.9)
    return c, H.Lattice.generate("triang", 96, c), np.geomspace(0.02, 2.0, 256)


def test_config3_fp32_256_replicas_energies_and_unit_spins():
    """The benchmarked FP32 engine (float4 LDS sites, 1024-site colour
    classes): after sweeps with the exchange, every replica's reported FP64
    energy equals the bond-list energy of the spins it holds and every spin
    stays a unit vector."""
    _say("config3 fp32 energies")
    c, lat, temps = _config3()
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=5, n_hb=1, n_or_boost=5,
                      n_hb_boost=1, pt_cap50=False, seed=3)
    rng = np.random.default_rng(303)
    mc.set_spins(_unit(rng, (256, lat.N, 3)))
    for n in (1, 4):
        mc.run(n, pt_parity=2)
        sp = mc.get_spins()
        np.testing.assert_allclose(np.linalg.norm(sp, axis=-1), 1.0, atol=1e-5)
        np.testing.assert_allclose(mc.energies(), lat.energies(sp, H.BATCHED_F32), rtol=1e-12,
                                   atol=1e-9)
    att, acc = mc.pt_counters()
    assert acc.sum() > 0


def test_config3_fp32_overrelaxation_sweep_matches_fp64_replay():
    """One overrelaxation pass (deterministic) over the nine 1024-site colour
    classes of all 256 FP32 replicas at full size against the oracle's FP64
    colour replay: pins the kernel's gathers, label packing and colour order
    at configuration 3 to FP32 rounding (2e-5).  A single pass: sequential
    overrelaxation amplifies rounding differences exponentially, and five
    passes (45 classes in sequence) take FP32 and FP64 apart completely."""
    _say("config3 fp32 OR replay")
    c, lat, temps = _config3()
    m = _model(lat, c)
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=1, n_hb=0, n_or_boost=1,
                      n_hb_boost=0, seed=3)
    rng = np.random.default_rng(304)
    s = _unit(rng, (256, lat.N, 3))
    mc.set_spins(s)
    mc.run(1, pt_parity=-1)
    ref = s.copy()
    m.mc_replay(0, temps, lat.colours(), ref, 1, n_or=1, n_hb=0, n_or_b=1, n_hb_b=0, seed=3)
    d = np.abs(mc.get_spins() - ref).max(axis=-1)
    err, q = d.max(), np.quantile(d, 0.9999)
    _say("max |dS| %.2e, 99.99 %% quantile %.2e" % (err, q))
    # the largest differences sit on the rare sites whose field nearly
    # cancels (the reflection divides by |h|^2)
    assert q < 2e-5
    assert err < 1e-3


def test_config3_fp32_heat_bath_sweep_matches_its_fp64_restatement():
    """One heat-bath pass (no overrelaxation) of all 256 FP32 replicas at
    full size against the oracle's FP64 evaluation of the FP32 kernel's own
    heat bath (its uniforms, the Philox block shared by a thread's sites of
    an even and the next odd colour class, its inverse CDF and frame): pins
    the kernel's random-stream indexing and heat-bath arithmetic at
    configuration 3.  The heat bath contracts differences, so FP32 rounding
    stays near 1e-6."""
    _say("config3 fp32 HB replay")
    c, lat, temps = _config3()
    m = _model(lat, c)
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=0, n_hb=1, n_or_boost=0,
                      n_hb_boost=1, seed=3)
    rng = np.random.default_rng(309)
    s = _unit(rng, (256, lat.N, 3))
    mc.set_spins(s)
    mc.run(1, pt_parity=-1)
    col = lat.colours()
    order = np.argsort(col, kind="stable").astype(np.int32)   # uniform degree: site order per class
    col_off = np.concatenate([[0], np.cumsum(np.bincount(col))]).astype(np.int32)
    ref = s.copy()
    m.mc_replay_hb32(temps, order, col_off, ref, n_hb=1, seed=3, sweep=0)
    d = np.abs(mc.get_spins() - ref).max(axis=-1)
    err, q = d.max(), np.quantile(d, 0.9999)
    far = (d > 1e-3).mean()
    _say("max |dS| %.2e, 99.99 %% quantile %.2e, fraction above 1e-3 %.1e" % (err, q, far))
    # the frame about h changes sign with h_z: where h_z is within rounding of
    # 0 the two precisions may pick opposite frames (one site in ~1e6)
    assert q < 1e-4
    assert far < 1e-5


@pytest.mark.parametrize("prec", [H.BATCHED_F64, H.PLAIN_F64])
def test_config3_fp64_replay_of_four_replicas(prec):
    """Configuration 3's lattice and couplings in FP64 (a replica of
    221 KB: the global-memory sweep), four replicas, three sweeps of
    5 OR + 1 HB with the exchange, each against the oracle's colour replay."""
    _say("config3 fp64 replay", prec)
    c, lat, _ = _config3()
    temps = np.array([0.05, 0.3, 0.9, 2.0])
    m = _model(lat, c)
    mc = H.MonteCarlo(lat, temps, precision=prec, n_or=5, n_hb=1, seed=13)
    rng = np.random.default_rng(305)
    s = _unit(rng, (4, lat.N, 3))
    mc.set_spins(s)
    worst, _ = _stepwise_replay(mc, m, lat, temps, s, 3, 1e-8, 5, 1, 2, 8, seed=13)
    _say("max |dS| %.2e" % worst)


def test_config3_ising_256_replicas_bit_exact():
    """The Ising variant (int8 spins, h = 0.3 on the Ising axis) of all 256
    replicas: five sweeps with the exchange, bit-exact against the oracle's
    colour replay (energy trace, slot map, counters, spins)."""
    _say("config3 ising")
    c, lat, temps = _config3()
    m = _model(lat, c)
    rng = np.random.default_rng(306)
    s = np.zeros((256, lat.N, 3))
    s[:, :, 0] = rng.choice([-1.0, 1.0], size=(256, lat.N))
    mc = H.MonteCarlo(lat, temps, model=H.ISING, precision=H.BATCHED_F32, seed=23)
    mc.set_spins(s)
    np.testing.assert_array_equal(mc.energies(), m.energies_ising(s))
    tr = mc.run(5, pt_parity=2, trace=True)
    ref = s.copy()
    rtr, (cos, cnt, exc, _) = m.mc_replay(1, temps, lat.colours(), ref, 5, seed=23, pt_parity=2)
    np.testing.assert_array_equal(tr, rtr)
    np.testing.assert_array_equal(mc.slot_map(), cos)
    att, acc = mc.pt_counters()
    np.testing.assert_array_equal(att, cnt)
    np.testing.assert_array_equal(acc, exc)
    np.testing.assert_array_equal(mc.get_spins(), ref[cos])
    assert acc.sum() > 0


# ------------------------------------------------------------ config 4

def _config4_lattices(rng):
    c = H.Couplings(J1=1.0, J2=0.6)
    lats = [H.Lattice.generate("square", 64, c), H.Lattice.generate("kagome", 32, c),
            H.Lattice.generate("hexagn", 40, c), H.Lattice.generate("square", 48, c)]
    vac = np.zeros(lats[3].N, np.int32)
    vac[rng.choice(lats[3].N, 230, replace=False)] = 1
    lats[3].set_vacancies(vac)
    return c, lats, vac


@pytest.mark.parametrize("form", [2, 1, 3])
def test_config4_fp32_mixed_batch_matches_oracle(form):
    """The four configuration-4 lattices in one FP32 batch (two samples
    each): four sites per thread (N up to 4096), default form 2 with each
    workgroup at its lattice's slot class (8 or 12: smd_kernel_two), form 1
    with the batch's largest class, form 3 in global memory.  30 RK4 steps
    (dt = 0.005) against the oracle: spins within 2e-6 (measured 3.7e-7),
    energies within 2e-7 relative."""
    _say("config4 fp32 batch form", form)
    rng = np.random.default_rng(404)
    c, lats, vac = _config4_lattices(rng)
    idx = np.array([0, 1, 2, 3, 3, 2, 1, 0], np.int32)
    samples = [_stripe_noise(rng, lats[i]) if i in (0, 3) else _unit(rng, (lats[i].N, 3))
               for i in idx]
    T = 31
    smd = H.SpinDynamics(lats, idx, 0.005, T, n_omega=32, precision=H.BATCHED_F32, dssf=False)
    smd.set_shape(form, 1)
    smd.set_samples(samples)
    smd.step(T - 1)
    out = smd.get_samples()
    E = smd.energies()
    models = [_model(lats[i], c, vac if i == 3 else None) for i in range(4)]
    worst = 0.0
    for k, i in enumerate(idx):
        s_ref, _, _, e = models[i].smd_run(samples[k], 0.005, T, gauss_seidel=False, traj=False)
        worst = max(worst, np.abs(out[k] - s_ref).max())
        # FP32 state: the FP64 energy of the FP32 spins drifts by rounding
        e_true = lats[i].energies(out[k][None], H.BATCHED_F64)[0]
        _say("sample %d: E engine %.6f, E(spins) %.6f, oracle %.6f, E(0) %.6f"
             % (k, E[k], e_true, e[-1], e[0]))
        assert E[k] == pytest.approx(e[-1], rel=2e-7)
        assert e_true == pytest.approx(e[-1], rel=2e-7)
    _say("max |dS| %.2e" % worst)
    assert worst < 2e-6


def _config4_dssf_case():
    rng = np.random.default_rng(406)
    c, lats, vac = _config4_lattices(rng)
    sub = [lats[3], lats[1]]
    idx = np.array([0, 1, 0, 1], np.int32)
    samples = [_stripe_noise(rng, sub[0]), _unit(rng, (sub[1].N, 3)), _stripe_noise(rng, sub[0]),
               _unit(rng, (sub[1].N, 3))]
    return c, sub, vac, idx, samples


def _feed_blocks(smd, trajs, step=256):
    T = trajs[0].shape[0]
    for t0 in range(0, T, step):
        t1 = min(T, t0 + step)
        smd.feed(np.concatenate([tr[t0:t1] for tr in trajs], axis=1))


@pytest.mark.parametrize("gemm,max_chunk", [(5, 2048), (5, None), (4, 2048), (4, None),
                                            (0, 2048), (2, 2048), (3, 2048), (1, 2048)])
def test_config4_mfma_dssf_of_a_known_trajectory_against_fp64(gemm, max_chunk, monkeypatch):
    """The split-FP16 MFMA contraction's own error: the oracle's FP64
    trajectories (square L=48 with its 230 vacancies and kagome L=32, two
    samples each, 2381 time points: with the chunk capped at 2048 points,
    one full chunk and a 333-point tail that is not a multiple of the
    contraction's 32-point step; uncapped, one 2381-point chunk) fed to the
    FP32 engine with each contraction tile, 512 bins, against the FP64 DSSF
    of the same trajectories (FFT form of the oracle's transform).  Bars:
    1e-4 of the peak, and 1e-3 relative on every bin above 1e-2 of its
    (q, component) peak; the FP64 engine fed the same points stays within
    1e-9."""
    if max_chunk:
        monkeypatch.setenv("HSMD_SMD_MAX_CHUNK", str(max_chunk))
    _say("config4 mfma dssf of a fed trajectory, gemm", gemm, "max chunk", max_chunk)
    c, sub, vac, idx, samples = _config4_dssf_case()
    T, no, dt = 2381, 512, 0.002
    models = [_model(sub[0], c, vac), _model(sub[1], c)]
    trajs = _pmap(lambda k: models[idx[k]].smd_run(samples[k], dt, T, gauss_seidel=False)[1],
                  range(4), workers=4)
    gms = []
    for li, L in enumerate(sub):
        g = L.grid_map().copy()
        if li == 0:
            g[(g >= 0) & (vac[np.maximum(g, 0)] != 0)] = -1   # vacant sites leave the DSSF
        gms.append(g)
    w = np.ones(T)
    ref = [sum(cpu_ref.dssf_fft(trajs[k], gms[li], sub[li].Gsz, w, no)
               for k in range(4) if idx[k] == li) for li in range(2)]
    for prec in (H.BATCHED_F32, H.BATCHED_F64):
        if prec == H.BATCHED_F64 and gemm != 0:
            continue
        smd = H.SpinDynamics(sub, idx, dt, T, n_omega=no, precision=prec)
        if prec == H.BATCHED_F32:
            smd.set_shape(0, gemm)
        _feed_blocks(smd, trajs)
        for li in range(2):
            d, _ = smd.dssf(li)
            a = ref[li]
            err = np.abs(d - a).max() / a.max()
            pk = a.max(axis=2, keepdims=True)
            big = a > 1e-2 * pk
            rel = (np.abs(d - a)[big] / a[big]).max()
            _say("prec %d lattice %d: max err / peak %.2e, max rel err above 1e-2 peak %.2e"
                 % (prec, li, err, rel))
            if prec == H.BATCHED_F32:
                assert err < 1e-4
                assert rel < 1e-3
            else:
                assert err < 1e-9


@pytest.mark.parametrize("max_chunk", [2048, None])
def test_pipelined_contraction_equals_the_128x512_tile(max_chunk, monkeypatch):
    """Contraction tile 5 (the pipelined 128 x 512 tile: twiddles in registers)
    forms the same split products in the same order as tile 4, so the DSSF
    of the same fed trajectory must be bit-identical: square L=48 with its
    vacancies and kagome L=32, two samples each, 2381 points of unit
    vectors, 512 bins, with the chunk capped at 2048 points (a chunk that
    stores Y, then a 333-point tail that adds to it and ends inside a
    32-point step) and uncapped (one 2381-point chunk)."""
    if max_chunk:
        monkeypatch.setenv("HSMD_SMD_MAX_CHUNK", str(max_chunk))
    _say("pipelined contraction vs 128 x 512 tile, max chunk", max_chunk)
    c, sub, vac, idx, samples = _config4_dssf_case()
    T, no, dt = 2381, 512, 0.002
    rng = np.random.default_rng(77)
    trajs = [_unit(rng, (T, sub[idx[k]].N, 3)) for k in range(4)]
    out = {}
    for gemm in (4, 5):
        smd = H.SpinDynamics(sub, idx, dt, T, n_omega=no, precision=H.BATCHED_F32)
        smd.set_shape(0, gemm)
        _feed_blocks(smd, trajs)
        out[gemm] = [smd.dssf(li)[0] for li in range(2)]
    for li in range(2):
        assert np.isfinite(out[5][li]).all()
        assert np.array_equal(out[4][li], out[5][li]), np.abs(out[4][li] - out[5][li]).max()


@pytest.mark.parametrize("no,max_chunk", [(512, None), (512, 1024), (200, 1024), (200, None)])
def test_fft_form_of_a_known_trajectory_against_fp64(no, max_chunk, monkeypatch):
    """Contraction form 6, the time DFT as a pruned FFT (T / 512 decimated
    512-point FFTs per row pair, combined per bin): square L=48 with its 230
    vacancies, kagome L=32 and square L=10 (300 rows: a ragged last block of
    32 rows), 2560 time points of unit vectors under a Gaussian window, fed
    to the FP32 engine, against the FP64 DSSF of the same trajectories (FFT
    form of the oracle's transform).  Uncapped the run is one chunk; capped
    at 1024 points it is two chunks that store and add and a 512-point tail,
    each decimated row cut by the chunk bounds.  At 200 bins the table holds
    256 (bins 200..255 zero).  Bars: 1e-5 of the peak, 1e-4 relative on
    every bin above 1e-2 of its (q, component) peak."""
    if max_chunk:
        monkeypatch.setenv("HSMD_SMD_MAX_CHUNK", str(max_chunk))
    _say("fft form dssf of a fed trajectory, n_omega", no, "max chunk", max_chunk)
    rng = np.random.default_rng(4061)
    c, lats, vac = _config4_lattices(rng)
    sub = [lats[3], lats[1], H.Lattice.generate("square", 10, c)]
    idx = np.array([0, 1, 2, 0, 2], np.int32)
    T, dt = 2560, 0.002
    tt = np.arange(T)
    w = np.exp(-((tt - T / 2) / (0.35 * T)) ** 2)
    trajs = [_unit(rng, (T, sub[i].N, 3)) for i in idx]
    gms = []
    for li, L in enumerate(sub):
        g = L.grid_map().copy()
        if li == 0:
            g[(g >= 0) & (vac[np.maximum(g, 0)] != 0)] = -1
        gms.append(g)
    ref = [sum(cpu_ref.dssf_fft(trajs[k], gms[li], sub[li].Gsz, w, no)
               for k in range(len(idx)) if idx[k] == li) for li in range(len(sub))]
    smd = H.SpinDynamics(sub, idx, dt, T, n_omega=no, precision=H.BATCHED_F32, window=w)
    smd.set_shape(0, 6)
    _feed_blocks(smd, trajs)
    for li in range(len(sub)):
        d, _ = smd.dssf(li)
        a = ref[li]
        assert d.shape == a.shape
        err = np.abs(d - a).max() / a.max()
        pk = a.max(axis=2, keepdims=True)
        big = a > 1e-2 * pk
        rel = (np.abs(d - a)[big] / a[big]).max()
        _say("lattice %d: max err / peak %.2e, max rel err above 1e-2 peak %.2e" % (li, err, rel))
        assert err < 1e-5
        assert rel < 1e-4


@pytest.mark.parametrize("T,no", [(512, 512), (1536, 384)])
def test_fft_form_short_runs_against_fp64(T, no):
    """The FFT form at P = T / 512 = 1 (the transform is one 512-point FFT per
    row pair, no combine over t2) and at an odd P = 3 with 384 bins: square
    L=10 (300 rows) and kagome L=8 (576 rows), three samples, fed unit
    vectors, Hann window, against the FP64 transform: within 1e-5 of the
    peak and 1e-4 relative above 1e-2 of the (q, component) peak."""
    _say("fft form short runs, T", T, "bins", no)
    rng = np.random.default_rng(4062 + T)
    c = H.Couplings(J1=1.0, J2=0.6)
    sub = [H.Lattice.generate("square", 10, c), H.Lattice.generate("kagome", 8, c)]
    idx = np.array([0, 1, 1], np.int32)
    w = 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(T) / T)
    trajs = [_unit(rng, (T, sub[i].N, 3)) for i in idx]
    ref = [sum(cpu_ref.dssf_fft(trajs[k], sub[li].grid_map(), sub[li].Gsz, w, no)
               for k in range(len(idx)) if idx[k] == li) for li in range(len(sub))]
    smd = H.SpinDynamics(sub, idx, 0.01, T, n_omega=no, precision=H.BATCHED_F32, window=w)
    smd.set_shape(0, 6)
    _feed_blocks(smd, trajs)
    for li in range(len(sub)):
        d, _ = smd.dssf(li)
        a = ref[li]
        err = np.abs(d - a).max() / a.max()
        big = a > 1e-2 * a.max(axis=2, keepdims=True)
        rel = (np.abs(d - a)[big] / a[big]).max()
        _say("lattice %d: max err / peak %.2e, max rel err above 1e-2 peak %.2e" % (li, err, rel))
        assert err < 1e-5
        assert rel < 1e-4


@pytest.mark.parametrize("gemm", [5, 4, 0, 2, 3, 1])
def test_config4_dssf_across_chunks_fp32_vs_fp64(gemm, monkeypatch):
    """The same batch integrated by both engines: FP32 RK4 with the MFMA
    contraction against FP64 RK4 with the FP64 GEMM over 2381 time points
    (dt = 0.002 keeps the two precisions' trajectories within FP32 rounding
    of each other over the run): max |a - b| <= 1e-3 of the peak, 1e-3
    relative on every bin above 1e-2 of its (q, component) peak (measured
    2.3e-6 and 1e-4), and qSS within 1e-3 of its peak."""
    monkeypatch.setenv("HSMD_SMD_MAX_CHUNK", "2048")   # a 2048-point chunk and a 333-point tail
    _say("config4 dssf chunks gemm", gemm)
    c, sub, vac, idx, samples = _config4_dssf_case()
    T, no, dt = 2381, 512, 0.002
    out = {}
    for prec in (H.BATCHED_F64, H.BATCHED_F32):
        smd = H.SpinDynamics(sub, idx, dt, T, n_omega=no, precision=prec)
        if prec == H.BATCHED_F32:
            smd.set_shape(0, gemm)
        smd.set_samples(samples)
        smd.step(1000)
        smd.step(T - 1 - 1000)
        out[prec] = [smd.dssf(li) for li in range(2)]
    for li in range(2):
        (a, qa), (b, qb) = out[H.BATCHED_F64][li], out[H.BATCHED_F32][li]
        err = np.abs(a - b).max() / a.max()
        pk = a.max(axis=2, keepdims=True)
        big = a > 1e-2 * pk
        rel = (np.abs(a - b)[big] / a[big]).max()
        _say("lattice %d: max err / peak %.2e, max rel err above 1e-2 peak %.2e (trajectory "
             "divergence included)" % (li, err, rel))
        assert err < 1e-3
        assert rel < 1e-3
        dq = np.abs(qb - qa).max() / qa.max()
        _say("lattice %d: qss max err / peak %.2e" % (li, dq))
        assert dq < 1e-3


@pytest.mark.parametrize("prec,tol", [(H.BATCHED_F32, 3e-6), (H.BATCHED_F64, 1e-10)])
def test_two_sites_per_thread_lds_form_matches_oracle(prec, tol):
    """Square L=40 (1600 sites: two sites per thread in the LDS-resident
    form, FP32 and FP64): 159 RK4 steps and the DSSF at 128 bins against the
    oracle: FP32 spins within 3e-6 (measured 5.7e-7), FP64 within 1e-10."""
    _say("spt=2", prec)
    c = H.Couplings(J1=1.0, J2=0.6, JX=0.1, h=0.2)
    lat = H.Lattice.generate("square", 40, c)
    m = _model(lat, c)
    rng = np.random.default_rng(407)
    samples = [_stripe_noise(rng, lat) for _ in range(3)]
    T, no, dt = 160, 128, 0.005
    w = cpu_ref.window("hann2", T)
    smd = H.SpinDynamics([lat], np.zeros(3, np.int32), dt, T, n_omega=no, precision=prec, window=w)
    smd.set_samples(samples)
    smd.step(T - 1)
    out = smd.get_samples()
    d, q = smd.dssf(0)
    ref = np.zeros_like(d)
    for k, s0 in enumerate(samples):
        s_ref, tr, _, _ = m.smd_run(s0, dt, T, gauss_seidel=False)
        err = np.abs(out[k] - s_ref).max()
        _say("sample %d max |dS| %.2e" % (k, err))
        assert err < tol, k
        ref += cpu_ref.dssf_fft(tr, lat.grid_map(), lat.Gsz, w, no)
    err = np.abs(d - ref).max() / ref.max()
    _say("dssf err / peak %.2e" % err)
    assert err < (1e-5 if prec == H.BATCHED_F32 else 1e-9)


# ------------------------------------------------------------ config 5

def _graph(n_target=12000):
    box = math.sqrt(n_target / 0.627)
    pts = H.poisson_disk(box, 1.0, seed=5)
    c = H.Couplings(J1=1.0, J2=0.4, JX=0.1)
    lat = H.Lattice.from_points(pts, box, (1.15, 1.45, 1.75), c)
    rng = np.random.default_rng(505)
    vac = np.zeros(lat.N, np.int32)
    vac[rng.choice(lat.N, int(round(0.05 * lat.N)), replace=False)] = 1
    lat.set_vacancies(vac)
    return c, lat, vac


def _q_path(n=200, qmax=2 * math.pi):
    corners = np.array([[0.0, 0.0], [qmax, 0.0], [qmax, qmax], [0.0, 0.0]])
    seg = np.linalg.norm(np.diff(corners, axis=0), axis=1)
    s = np.linspace(0.0, seg.sum(), n)
    out = np.zeros((n, 2))
    cs = np.cumsum(seg)
    for k, v in enumerate(s):
        i = min(np.searchsorted(cs, v, side="right"), 2)
        t = (v - (cs[i - 1] if i > 0 else 0.0)) / seg[i]
        out[k] = corners[i] + min(max(t, 0.0), 1.0) * (corners[i + 1] - corners[i])
    return out


def test_config5_graph_fp32_128_replicas_tuned_overrelaxation_replay():
    """The Poisson-disk graph (N ~ 12,000, 5 % vacancies) with the
    autotuned colour schedule and workgroup width: one overrelaxation pass
    of all 128 FP32 replicas (compact labels) against the oracle's FP64
    replay of that schedule, then sweeps with heat bath and exchange whose
    energies equal the bond-list energies."""
    _say("config5 mc fp32")
    c, lat, vac = _graph()
    assert 11000 < lat.N < 13000
    m = _model(lat, c, vac)
    temps = np.geomspace(0.05, 1.5, 128)
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=1, n_hb=0, n_or_boost=1,
                      n_hb_boost=0, seed=5, tune=True, tune_sweeps=3)
    _say("tuning", mc.tuning)
    rng = np.random.default_rng(506)
    s = _unit(rng, (128, lat.N, 3))
    mc.set_spins(s)
    mc.run(1, pt_parity=-1)
    ref = s.copy()
    m.mc_replay(0, temps, lat.colours(), ref, 1, n_or=1, n_hb=0, n_or_b=1, n_hb_b=0, seed=5)
    got = mc.get_spins()
    live = vac == 0
    d = np.abs(got[:, live] - ref[:, live]).max(axis=-1)
    err, q = d.max(), np.quantile(d, 0.9999)
    _say("max |dS| %.2e, 99.99 %% quantile %.2e" % (err, q))
    assert q < 2e-5
    assert err < 1e-3
    mh = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=5, n_hb=1, n_or_boost=5,
                      n_hb_boost=1, seed=6)
    mh.set_spins(s)
    mh.run(3, pt_parity=2)
    sp = mh.get_spins()
    np.testing.assert_allclose(np.linalg.norm(sp[:, live], axis=-1), 1.0, atol=1e-5)
    np.testing.assert_allclose(mh.energies(), lat.energies(sp, H.BATCHED_F32), rtol=1e-12,
                               atol=1e-9)


@pytest.mark.parametrize("prec", [H.BATCHED_F64, H.PLAIN_F64])
def test_config5_graph_fp64_replay(prec):
    """The Poisson-disk graph in FP64, four replicas, three sweeps of
    5 OR + 1 HB with the exchange, each against the oracle's colour replay
    (vacant sites carry no couplings: their spins are updated but not
    compared)."""
    _say("config5 mc fp64", prec)
    c, lat, vac = _graph()
    m = _model(lat, c, vac)
    temps = np.array([0.05, 0.2, 0.6, 1.5])
    mc = H.MonteCarlo(lat, temps, precision=prec, n_or=5, n_hb=1, seed=7)
    rng = np.random.default_rng(507)
    s = _unit(rng, (4, lat.N, 3))
    mc.set_spins(s)
    worst, _ = _stepwise_replay(mc, m, lat, temps, s, 3, 1e-8, 5, 1, 2, 8, seed=7)
    _say("max |dS| %.2e" % worst)


@pytest.mark.parametrize("nq", [200, 300])
def test_config5_qpath_mfma_dssf_of_a_known_trajectory(nq):
    """The FP32 engine's split-FP16 MFMA q-path (sites as the reduction
    index, wavevectors in chunks of 256) on the oracle's own FP64
    trajectories of the graph (four samples, 256 time points, 5 % vacancies)
    fed to the engine, against the oracle's direct sums along the path
    (200 points; 300 spans two wavevector chunks).  Bars: 1e-4 of the peak
    and 1e-3 relative on every bin above 1e-2 of its (q, component) peak;
    the FP64 engine fed the same points within 1e-9."""
    _say("config5 q-path mfma of a fed trajectory, nq", nq)
    c, lat, vac = _graph()
    m = _model(lat, c, vac)
    rng = np.random.default_rng(509)
    S, T, no, dt = 4, 256, 128, 0.01
    samples = [_unit(rng, (lat.N, 3)) for _ in range(S)]
    trajs = _pmap(lambda s0: m.smd_run(s0, dt, T, gauss_seidel=False)[1], samples, workers=4)
    q = _q_path(nq, qmax=3 * math.pi)
    rv = lat.positions()
    ref = sum(cpu_ref.dssf_direct_fft(tr, rv, q, np.ones(T), no, mask=vac) for tr in trajs)
    for prec in (H.BATCHED_F32, H.BATCHED_F64):
        smd = H.SpinDynamics([lat], np.zeros(S, np.int32), dt, T, n_omega=no, precision=prec)
        _feed_blocks(smd, trajs, step=64)
        d = smd.dssf_q(q)
        err = np.abs(d - ref).max() / ref.max()
        pk = ref.max(axis=2, keepdims=True)
        big = ref > 1e-2 * pk
        rel = (np.abs(d - ref)[big] / ref[big]).max()
        _say("prec %d: max err / peak %.2e, max rel err above 1e-2 peak %.2e" % (prec, err, rel))
        if prec == H.BATCHED_F32:
            assert err < 1e-4
            assert rel < 1e-3
        else:
            assert err < 1e-9


@pytest.mark.parametrize("prec,tol", [(H.BATCHED_F32, 3e-5), (H.BATCHED_F64, 1e-10)])
def test_config5_graph_smd_and_qpath_dssf_match_oracle(prec, tol):
    """SMD on the graph in the global-memory form (autotuned shape for
    FP32): four samples, 255 RK4 steps (dt = 0.01) against the oracle; then
    the DSSF along the 200-point q-path by direct spatial sums against the
    oracle's (vacant sites left out)."""
    _say("config5 smd + q-path", prec)
    c, lat, vac = _graph()
    m = _model(lat, c, vac)
    rng = np.random.default_rng(508)
    S, T, no, dt = 4, 256, 128, 0.01
    samples = [_unit(rng, (lat.N, 3)) for _ in range(S)]
    smd = H.SpinDynamics([lat], np.zeros(S, np.int32), dt, T, n_omega=no, precision=prec,
                         tune=(prec == H.BATCHED_F32), tune_steps=16)
    smd.set_samples(samples)
    smd.step(T - 1)
    out = smd.get_samples()
    q = _q_path(200)
    d = smd.dssf_q(q)
    rv = lat.positions()
    live = vac == 0
    w = np.ones(T)

    def one(s0):
        s_ref, tr, _, _ = m.smd_run(s0, dt, T, gauss_seidel=False)
        return s_ref, cpu_ref.dssf_direct_fft(tr, rv, q, w, no, mask=vac)

    res = _pmap(one, samples, workers=4)
    for k in range(S):
        assert np.abs(out[k][live] - res[k][0][live]).max() < tol, k
    ref = sum(r[1] for r in res)
    err = np.abs(d - ref).max() / ref.max()
    _say("q-path dssf err / peak %.2e" % err)
    assert err < (1e-3 if prec == H.BATCHED_F32 else 1e-9)
