"""GPU parity tests of the MC engine through the C ABI, against the oracle."""
import math
import sys

import numpy as np
import pytest

import classical_mc_and_smd_for_2d_heisenberg_models_amd as H
import cpu_ref

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-mc]", *a, flush=True)
    sys.stdout.flush()


def _spins(rng, R, N):
    v = rng.normal(size=(R, N, 3))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _model(lat, c, geom=None, imp=None):
    return cpu_ref.Model(lat.tables(), (lat.zn1, lat.zn2, lat.znx), c.as_params(), imp=imp,
                         lieb=(geom == "lieb"))


def test_philox_and_exp_are_bit_identical_to_oracle():
    _say("philox/exp")
    seed = 0x1234_5678_9abc_def0
    g = H.debug_philox(seed, 77, 5, 1, 1000)
    for i in [0, 1, 17, 999]:
        np.testing.assert_array_equal(g[i], cpu_ref.philox(seed, [i, 77, 5, 1]))
    rng = np.random.default_rng(0)
    x = np.concatenate([-rng.random(5000) * 720, -rng.random(500) * 1e-3, [0.0, 2.0, -699.9]])
    y = H.debug_exp(x)
    ref = np.array([cpu_ref.det_exp(v) for v in x])
    np.testing.assert_array_equal(y.view(np.uint64), ref.view(np.uint64))


FIELD_CASES = [
    ("square", 24, dict(J1=1.0, J2=0.6), None),
    ("triang", 12, dict(J1=1.0, J2=0.12, JX=0.05, h=0.3, lambda1=0.9, lambda2=0.9), None),
    ("kagome", 6, dict(J1=1.0, J2=-0.2, JX=0.1, lambda1=1.2, lambda2=0.8, h=-0.4), None),
    ("hexagn", 6, dict(J1=0.7, J2=0.3, JX=0.2), None),
    ("lieb", 6, dict(J1=1.0, J2=0.2, J2_delta=0.05), None),
    ("square", 12, dict(J1=1.0, J2=0.5, imp_ratio=0.2, J1ic=0.5, J1ii=0.3, J2ic=0.1, J2ii=0.05), 0.2),
]


def test_poisson_disk_graph_fp32_compact_labels_energies_and_graph_loop():
    """FP32 sweeps on the sparse Poisson-disk table (compact labels): unit
    spins, energies equal to the bond-list sum of the spins left, and the
    replayed graph loop equal to the launch-by-launch loop."""
    _say("poisson graph fp32 compact")
    box = 40.0
    pts = H.poisson_disk(box, 1.0, seed=9)
    c = H.Couplings(J1=1.0, J2=0.4, JX=0.1, h=0.2)
    lat = H.Lattice.from_points(pts, box, (1.15, 1.45, 1.75), c)
    temps = np.geomspace(0.1, 2.0, 6)
    rng = np.random.default_rng(41)
    s = _spins(rng, 6, lat.N)
    a = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=3, n_hb=1, seed=17)
    b = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=3, n_hb=1, seed=17)
    a.set_spins(s)
    b.set_spins(s)
    a.run(6, pt_parity=2)                    # replayed graph
    for _ in range(6):
        b.run(1, pt_parity=2)                # launch by launch
    np.testing.assert_array_equal(a.get_spins(), b.get_spins())
    np.testing.assert_array_equal(a.energies(), b.energies())
    sp = a.get_spins()
    np.testing.assert_allclose(np.linalg.norm(sp, axis=-1), 1.0, atol=1e-5)
    np.testing.assert_allclose(a.energies(), lat.energies(sp, H.BATCHED_F32), rtol=1e-12, atol=1e-10)


SWEEP_ENERGY_CASES = [
    ("square", 24, dict(J1=1.0, J2=0.6), True),
    ("triang", 12, dict(J1=1.0, J2=0.12, JX=0.05, h=0.3, lambda1=0.9, lambda2=0.9), True),
    ("kagome", 6, dict(J1=1.0, J2=-0.2, JX=0.1, lambda1=1.2, lambda2=0.8, h=-0.4), True),
    ("hexagn", 6, dict(J1=0.7, J2=0.3, JX=0.2), True),
    ("square", 8, dict(J1=1.0, J2=0.4, h=0.2), False),
    ("triang", 6, dict(J1=1.0, J2=0.3), True),
]


@pytest.mark.parametrize("prec", [H.BATCHED_F32, H.BATCHED_F64])
@pytest.mark.parametrize("geom,L,kw,pbc", SWEEP_ENERGY_CASES)
def test_sweep_energies_equal_bond_list_energies(geom, L, kw, pbc, prec):
    """The energy each sweep kernel reports (table form or bond list, as the
    engine picks) equals get_energyValue's bond-list sum over the spins the
    sweep left, in FP64 (products of the stored spins), to 1e-12."""
    _say("sweep energy", geom, L, pbc, prec)
    c = H.Couplings(**kw)
    lat = H.Lattice.generate(geom, L, c, pbc=pbc)
    temps = np.array([0.2, 0.7, 1.4, 2.5, 4.0])
    rng = np.random.default_rng(31)
    mc = H.MonteCarlo(lat, temps, precision=prec, seed=4)
    mc.set_spins(_spins(rng, len(temps), lat.N))
    for n, parity in [(1, -1), (3, 2), (2, 0)]:
        mc.run(n, pt_parity=parity)
        E = mc.energies()
        ref = lat.energies(mc.get_spins(), prec)
        np.testing.assert_allclose(E, ref, rtol=1e-12, atol=1e-11)


@pytest.mark.parametrize("geom,L,kw,imp", FIELD_CASES)
def test_local_fields_and_energies_fp64_match_oracle(geom, L, kw, imp):
    _say("fields", geom, L, kw)
    c = H.Couplings(**kw)
    lat = H.Lattice.generate(geom, L, c)
    rng = np.random.default_rng(L)
    impmap = None
    if imp:
        impmap = (rng.random(lat.N) < imp).astype(np.int32)
        lat.set_impurities(impmap)
    m = _model(lat, c, geom, impmap)
    s = _spins(rng, 3, lat.N)
    for prec in (H.BATCHED_F64, H.PLAIN_F64):
        h = lat.local_fields(s, prec)
        np.testing.assert_allclose(h, m.local_fields(s), rtol=1e-12, atol=1e-13)
        E = lat.energies(s, prec)
        np.testing.assert_allclose(E, m.energies(s), rtol=1e-12, atol=1e-11)
    # FP32 state: fields to FP32 rounding, energies reduced in FP64
    h32 = lat.local_fields(s, H.BATCHED_F32)
    np.testing.assert_allclose(h32, m.local_fields(s), atol=2e-5)
    s32 = s.astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(lat.energies(s, H.BATCHED_F32), m.energies(s32), rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("prec", [H.BATCHED_F32, H.PLAIN_F64])
@pytest.mark.parametrize("case", ["triang_h", "square_disorder"])
def test_ising_metropolis_is_bit_exact_against_colour_replay(prec, case):
    _say("ising replay", prec, case)
    rng = np.random.default_rng(21)
    imp = None
    if case == "triang_h":
        geom, L, c = "triang", 12, H.Couplings(J1=1.0, J2=0.12, JX=0.05, h=0.3)
        lat = H.Lattice.generate(geom, L, c)
    else:
        geom, L, c = "square", 10, H.Couplings(J1=1.0, J2=0.4, imp_ratio=0.3, J1ic=0.6, J1ii=-0.2,
                                                J2ic=0.1, J2ii=0.3)
        lat = H.Lattice.generate(geom, L, c)
        imp = (rng.random(lat.N) < 0.3).astype(np.int32)
        lat.set_impurities(imp)
    m = _model(lat, c, geom, imp)
    temps = np.geomspace(0.5, 5.0, 8)
    R = len(temps)
    s = np.zeros((R, lat.N, 3))
    s[:, :, 0] = rng.choice([-1.0, 1.0], size=(R, lat.N))
    mc = H.MonteCarlo(lat, temps, model=H.ISING, precision=prec, seed=99)
    mc.set_spins(s)
    np.testing.assert_array_equal(mc.energies(), m.energies_ising(s))
    tr = mc.run(30, pt_parity=2, trace=True)
    ref = s.copy()
    rtr, (cos, cnt, exc, boost) = m.mc_replay(1, temps, lat.colours(), ref, 30, seed=99,
                                              pt_parity=2)
    np.testing.assert_array_equal(tr, rtr)
    np.testing.assert_array_equal(mc.slot_map(), cos)
    att, acc = mc.pt_counters()
    np.testing.assert_array_equal(att, cnt)
    np.testing.assert_array_equal(acc, exc)
    np.testing.assert_array_equal(mc.get_spins(), ref[cos])
    assert acc.sum() > 0


@pytest.mark.parametrize("prec", [H.BATCHED_F64, H.PLAIN_F64])
def test_heisenberg_fp64_sweep_matches_colour_replay(prec):
    _say("heis fp64 replay", prec)
    c = H.Couplings(J1=1.0, J2=0.6, h=0.2, lambda1=0.9, lambda2=1.1)
    lat = H.Lattice.generate("square", 8, c)
    m = _model(lat, c)
    temps = np.array([0.3, 0.7, 1.5])
    rng = np.random.default_rng(5)
    s = _spins(rng, 3, lat.N)
    mc = H.MonteCarlo(lat, temps, precision=prec, n_or=2, n_hb=2, seed=7)
    mc.set_spins(s)
    tr = mc.run(2, pt_parity=-1, trace=True)
    ref = s.copy()
    rtr, _ = m.mc_replay(0, temps, lat.colours(), ref, 2, n_or=2, n_hb=2, seed=7)
    np.testing.assert_allclose(mc.get_spins(), ref, atol=1e-9)
    np.testing.assert_allclose(tr, rtr, rtol=1e-10)


def test_heisenberg_fp32_open_chain_energy_matches_exact():
    # classical open chain: <S_i.S_i+1> = 1/(bJ) - coth(bJ) exactly; pins the
    # FP32 heat bath + overrelaxation statistically
    _say("open chain fp32")
    N = 512
    nb = -np.ones((N, 3), np.int32)
    for i in range(N):
        if i > 0: nb[i, 0] = i - 1
        if i < N - 1: nb[i, 1] = i + 1
    lat = H.Lattice.from_tables(nb, 2, 0, 1, None, H.Couplings(J1=1.0))
    temps = np.array([0.25, 0.5, 1.0, 2.0])
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=1, n_hb=1, seed=3)
    rng = np.random.default_rng(1)
    mc.set_spins(_spins(rng, len(temps), N))
    mc.run(500)
    tr = mc.run(4000, trace=True)
    for k, T in enumerate(temps):
        b = 1.0 / T
        exact = (N - 1) * (1 / b - 1 / math.tanh(b))
        est = tr[:, k].mean()
        # blocked error estimate
        blocks = tr[:, k].reshape(40, -1).mean(axis=1)
        err = blocks.std(ddof=1) / math.sqrt(len(blocks))
        assert abs(est - exact) < 5 * err + 1e-3 * abs(exact), (T, est, exact, err)


def test_heisenberg_fp32_matches_reference_order_statistics():
    # square J1-J2 at three temperatures: GPU FP32 colour sweeps vs the
    # oracle's reference-order FP64 chain, mean energies within error bars
    _say("square fp32 vs reference order")
    c = H.Couplings(J1=1.0, J2=0.6)
    lat = H.Lattice.generate("square", 8, c)
    m = _model(lat, c)
    temps = np.array([0.3, 0.6, 1.2])
    rng = np.random.default_rng(2)
    s = _spins(rng, 3, lat.N)
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=5, n_hb=1, seed=11)
    mc.set_spins(s)
    mc.run(1000)
    tr = mc.run(8000, trace=True)
    ref = s.copy()
    nref = 8000
    m.mc_reference(0, temps, ref, 1000, n_or=5, n_hb=1, pt_start=10 ** 9, seed=5)
    cos, acc = m.mc_reference(0, temps, ref, nref, n_or=5, n_hb=1, pt_start=10 ** 9, seed=6)
    for k in range(3):
        g = tr[:, k]
        gb = g.reshape(40, -1).mean(axis=1)
        ge = gb.std(ddof=1) / math.sqrt(40)
        rmean = acc[k, 0] / nref
        rvar = acc[k, 1] / nref - rmean ** 2
        re = math.sqrt(rvar / nref) * 4  # autocorrelation allowance
        assert abs(g.mean() - rmean) < 3 * math.hypot(ge, re) + 0.05, (temps[k], g.mean(), rmean)


def test_pt_moves_configurations_and_respects_cap():
    _say("pt counters")
    c = H.Couplings(J1=1.0, J2=0.6)
    lat = H.Lattice.generate("square", 8, c)
    temps = np.geomspace(0.05, 1.5, 16)
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=5, n_hb=1, seed=1, pt_cap50=True)
    rng = np.random.default_rng(0)
    mc.set_spins(_spins(rng, 16, lat.N))
    mc.run(300, pt_parity=2)
    att, acc = mc.pt_counters()
    assert att.min() > 0
    # the receiver's ratio (acc+1)/att stays within the 50 % cap (+1 slack)
    assert np.all(acc <= 0.5 * att + 2)
    perm = mc.slot_map()
    assert sorted(perm.tolist()) == list(range(16))
    # energies stay ordered on average with temperature
    E = mc.run(200, pt_parity=2, trace=True).mean(axis=0)
    assert np.all(np.diff(E) > -2.0)


@pytest.mark.parametrize("geom,L,kw,model", [
    ("square", 8, dict(J1=1.0, J2=0.6), H.HEISENBERG),
    ("triang", 6, dict(J1=1.0, h=0.4), H.HEISENBERG),
    ("lieb", 4, dict(J1=1.0, J2=0.3), H.HEISENBERG),
    ("square", 8, dict(J1=1.0, J2=0.6), H.ISING),
])
def test_measured_order_parameters_match_oracle(geom, L, kw, model):
    # one measured sweep: the per-slot moments equal the oracle's order
    # parameters (absStaggMag, absStripeMag, IsingZ2Parameter / cplxC3Parameter,
    # |M|, M_z) of the configuration after the sweep
    _say("observables", geom, L, model)
    c = H.Couplings(**kw)
    lat = H.Lattice.generate(geom, L, c)
    temps = np.array([0.3, 0.8, 1.5, 3.0])
    rng = np.random.default_rng(13)
    if model == H.ISING:
        s = np.zeros((4, lat.N, 3)); s[:, :, 0] = rng.choice([-1.0, 1.0], size=(4, lat.N))
    else:
        s = _spins(rng, 4, lat.N)
    mc = H.MonteCarlo(lat, temps, model=model, precision=H.BATCHED_F64, n_or=1, n_hb=1, seed=2)
    mc.set_spins(s)
    mc.run(3, pt_parity=2)
    mc.moments(reset=True)
    mc.run(1, pt_parity=-1, measure=True)
    mom, n = mc.moments()
    assert n == 1
    sp = mc.get_spins()
    E = mc.energies()
    nb = lat.tables()
    zn = (lat.zn1, lat.zn2, lat.znx)
    for k in range(4):
        o = cpu_ref.observables(geom, L, nb, zn, int(lat.flags["h_on"]), sp[k])
        F1, T1, Q1, SH, SV, SD, S1, X1 = o
        m = mom[k]
        np.testing.assert_allclose(m[0], E[k], rtol=1e-12)
        np.testing.assert_allclose([m[2], m[5], m[7], m[13], m[14], m[15], m[10]],
                                   [F1, T1, Q1, SH, SV, SD, S1], rtol=1e-10, atol=1e-12)
        x = abs(X1) if geom == "square" else X1
        np.testing.assert_allclose(m[16], x, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose([m[3], m[8], m[11], m[17]], [F1 ** 2, Q1 ** 2, S1 ** 2, X1 ** 2],
                                   rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("prec", [H.PLAIN_F64, H.BATCHED_F64])
def test_poisson_disk_graph_fp64_sweep_matches_colour_replay(prec):
    # MC on an irregular tabulated graph: one FP64 sweep replayed on the CPU
    # (the batched form reads the sparse table as compact labels)
    _say("poisson graph mc replay", prec)
    box = 30.0
    pts = H.poisson_disk(box, 1.0, seed=7)
    c = H.Couplings(J1=1.0, J2=0.4, JX=0.1)
    lat = H.Lattice.from_points(pts, box, (1.15, 1.45, 1.75), c)
    R = 4
    temps = np.geomspace(0.1, 1.0, R)
    mc = H.MonteCarlo(lat, temps, precision=prec, n_or=2, n_hb=1, seed=21)
    rng = np.random.default_rng(22)
    s = _spins(rng, R, lat.N)
    mc.set_spins(s)
    mc.run(1, pt_parity=-1)
    got = mc.get_spins()
    m = _model(lat, c)
    ref = s.copy()
    m.mc_replay(0, temps, lat.colours(), ref, 1, n_or=2, n_hb=1, seed=21)
    np.testing.assert_allclose(got, ref, atol=1e-9)


def _binned(x, nb):
    b = np.asarray(x, dtype=np.float64).reshape(nb, -1).mean(axis=1)
    return b.mean(), b.std(ddof=1) / math.sqrt(nb)


def test_observables_match_reference_order_chain_within_error_bars():
    # E, C, |M|, the stripe order parameter and the stripe S(q) peak of the
    # GPU FP32 colour-order chain against the oracle's reference-order FP64
    # chain (random-permutation sweeps, its own counter stream), square J1-J2
    # with J2 = 0.6 (stripe phase at low T), no exchange.  Every observable
    # agrees within 3 sigma and at least 80 % within 2 sigma.
    _say("observables vs reference order")
    c = H.Couplings(J1=1.0, J2=0.6)
    L = 8
    lat = H.Lattice.generate("square", L, c)
    N = lat.N
    m = _model(lat, c)
    temps = np.array([0.3, 0.6, 1.2])
    R, nbins, per_bin, every = 3, 20, 400, 10
    rng = np.random.default_rng(31)
    s0 = _spins(rng, R, N)
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=5, n_hb=1, seed=17)
    mc.set_spins(s0)
    mc.set_sq(every)
    mc.run(1000)
    gpu = {k: [] for k in ("E", "E2", "M", "S", "sq")}
    for b in range(nbins):
        mc.moments(reset=True)
        mc.run(per_bin, measure=True)
        mom, n = mc.moments()
        gpu["E"].append(mom[:, 0] / n / N)
        gpu["E2"].append(mom[:, 1] / n / N ** 2)
        gpu["M"].append(mom[:, 2] / n)
        gpu["S"].append(mom[:, 10] / n)
        # xy S(q) peaks at X and Y (Values2 order: xy block, points 1 and 2)
        gpu["sq"].append((mom[:, 26] + mom[:, 27]) / mom[:, 65])
    gpu = {k: np.array(v) for k, v in gpu.items()}
    # oracle chain, measured every `every` sweeps
    ref = s0.copy()
    m.mc_reference(0, temps, ref, 1000, n_or=5, n_hb=1, pt_start=10 ** 9, seed=1001)
    nb, rv, gm, zn = cpu_ref.lattice("square", L)
    cpu = {k: [] for k in ("E", "E2", "M", "S", "sq")}
    nmeas = nbins * per_bin // every
    for t in range(nmeas):
        m.mc_reference(0, temps, ref, every, n_or=5, n_hb=1, pt_start=10 ** 9, seed=2000 + t)
        E = m.energies(ref) / N
        sq = cpu_ref.sq_grid(gm, L, ref)
        row = {"E": E, "E2": E ** 2, "M": [], "S": [], "sq": []}
        for k in range(R):
            o = cpu_ref.observables("square", L, nb, zn, 0, ref[k])
            pk, _ = cpu_ref.corr_len(sq[k], L, False, 2 * math.pi / L)
            row["M"].append(o[0])
            row["S"].append(o[6])
            row["sq"].append(pk[0, 1] + pk[0, 2])
        for kk in cpu:
            cpu[kk].append(row[kk])
    cpu = {k: np.array(v) for k, v in cpu.items()}
    within2 = 0
    total = 0
    for k in range(R):
        beta = 1.0 / temps[k]
        for name in ("E", "M", "S", "sq", "C"):
            if name == "C":
                # specific heat per site from per-bin <e^2> - <e>^2, bin errors
                gc = N * beta ** 2 * (gpu["E2"][:, k] - gpu["E"][:, k] ** 2)
                cc = N * beta ** 2 * (cpu["E2"][:, k].reshape(nbins, -1).mean(1)
                                      - cpu["E"][:, k].reshape(nbins, -1).mean(1) ** 2)
                ga, ge = gc.mean(), gc.std(ddof=1) / math.sqrt(nbins)
                ca, ce = cc.mean(), cc.std(ddof=1) / math.sqrt(nbins)
            else:
                ga, ge = _binned(gpu[name][:, k], nbins)
                ca, ce = _binned(cpu[name][:, k], nbins)
            sig = math.hypot(ge, ce) + 1e-12
            _say("T=%.2f %-3s gpu %.5f +- %.5f  ref %.5f +- %.5f  (%.2f sigma)"
                 % (temps[k], name, ga, ge, ca, ce, abs(ga - ca) / sig))
            assert abs(ga - ca) < 3 * sig + 1e-9, (temps[k], name, ga, ca, sig)
            within2 += abs(ga - ca) < 2 * sig
            total += 1
    assert within2 >= 0.8 * total


@pytest.mark.parametrize("R", [1, 7, 1024])
def test_ising_replay_at_replica_count_edges(R):
    """One replica (no exchange pairs), an odd count (the last slot has no
    partner on one parity) and the largest count (1024, one exchange thread
    per pair, a full grid of sweep workgroups): bit-exact against the
    oracle's colour replay with alternating pair parity."""
    _say("ising replay R", R)
    c = H.Couplings(J1=1.0, J2=0.3)
    lat = H.Lattice.generate("square", 6 if R == 1024 else 10, c)
    m = _model(lat, c, "square")
    temps = np.geomspace(0.5, 5.0, R)
    rng = np.random.default_rng(R)
    s = np.zeros((R, lat.N, 3))
    s[:, :, 0] = rng.choice([-1.0, 1.0], size=(R, lat.N))
    mc = H.MonteCarlo(lat, temps, model=H.ISING, precision=H.BATCHED_F32, seed=5)
    mc.set_spins(s)
    n = 12
    tr = mc.run(n, pt_parity=2, trace=True)
    ref = s.copy()
    rtr, (cos, cnt, exc, boost) = m.mc_replay(1, temps, lat.colours(), ref, n, seed=5, pt_parity=2)
    np.testing.assert_array_equal(tr, rtr)
    np.testing.assert_array_equal(mc.slot_map(), cos)
    att, acc = mc.pt_counters()
    np.testing.assert_array_equal(att, cnt)
    np.testing.assert_array_equal(acc, exc)
    np.testing.assert_array_equal(mc.get_spins(), ref[cos])
    if R == 1:
        assert att.sum() == 0
    else:
        assert acc.sum() > 0
    # replayed graph form of the same chain
    mc.run(6, pt_parity=2)
    m.mc_replay(1, temps, lat.colours(), ref, 6, seed=5, sweep0=n, pt_parity=2,
                state=(cos, cnt, exc, boost))
    np.testing.assert_array_equal(mc.get_spins(), ref[cos])


def test_replica_count_limits():
    _say("replica limits")
    lat = H.Lattice.generate("square", 4, H.Couplings(J1=1.0))
    with pytest.raises(H.HSMDError):
        H.MonteCarlo(lat, np.linspace(1.0, 2.0, 1025))
    with pytest.raises(H.HSMDError):
        H.MonteCarlo(lat, np.zeros(0))


def test_heisenberg_fp64_beyond_lds_matches_colour_replay():
    """A replica too large for LDS in FP64 (square L=96: 3 x 9217 x 8 B =
    221 KB) runs the global-memory sweep; it matches the oracle's FP64
    colour replay as the LDS-resident form does."""
    _say("heis fp64 beyond LDS")
    c = H.Couplings(J1=1.0, J2=0.6, h=0.2, lambda1=0.9, lambda2=1.1)
    lat = H.Lattice.generate("square", 96, c)
    m = _model(lat, c)
    temps = np.array([0.3, 1.5])
    rng = np.random.default_rng(8)
    s = _spins(rng, 2, lat.N)
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F64, n_or=2, n_hb=2, seed=11)
    mc.set_spins(s)
    tr = mc.run(2, pt_parity=-1, trace=True)
    ref = s.copy()
    rtr, _ = m.mc_replay(0, temps, lat.colours(), ref, 2, n_or=2, n_hb=2, seed=11)
    # 18,432 heat-bath updates per replica use the device's and the host's
    # FP64 log/exp/sincos (each within an ulp or two); a site whose draw
    # falls on a steep part of the inverse CDF can carry that to ~1e-9
    np.testing.assert_allclose(mc.get_spins(), ref, atol=1e-8)
    np.testing.assert_allclose(tr, rtr, rtol=1e-10)


def test_fp32_beyond_lds_energies_and_unit_spins():
    """FP32 replicas too large for LDS (square L=128, 197 KB) sweep in global
    memory: unit spins, and energies equal to the bond-list sum of the spins."""
    _say("fp32 beyond LDS")
    c = H.Couplings(J1=1.0, J2=0.5, JX=0.1, h=0.1)
    lat = H.Lattice.generate("square", 128, c)
    temps = np.array([0.4, 1.0, 3.0])
    rng = np.random.default_rng(9)
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, seed=12)
    mc.set_spins(_spins(rng, 3, lat.N))
    mc.run(4, pt_parity=2)
    sp = mc.get_spins()
    np.testing.assert_allclose(np.linalg.norm(sp, axis=-1), 1.0, atol=1e-5)
    np.testing.assert_allclose(mc.energies(), lat.energies(sp, H.BATCHED_F32), rtol=1e-12, atol=1e-10)
