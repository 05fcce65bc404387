"""The autotuners (MC colour schedule x workgroup shape; SMD workgroup form x
DSSF contraction tile): their choices are valid, cached, and engines run
with them keep parity -- the Ising chain with a tuned colour schedule is
still bit-exact against the oracle's colour replay of that schedule."""
import sys

import numpy as np
import pytest

import classical_mc_and_smd_for_2d_heisenberg_models_amd as H
import cpu_ref

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-tune]", *a, flush=True)
    sys.stdout.flush()


def _spins(rng, R, N):
    v = rng.normal(size=(R, N, 3))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def test_mc_tuner_choice_is_valid_and_cached(tmp_path, monkeypatch):
    _say("mc tuner")
    monkeypatch.setenv("HSMD_TUNE_CACHE", str(tmp_path / "tune.txt"))
    rng = np.random.default_rng(2)
    pts = H.poisson_disk(30.0, 1.0, seed=4)
    c = H.Couplings(J1=1.0, J2=0.4, JX=0.1)
    lat = H.Lattice.from_points(pts, 30.0, (1.15, 1.45, 1.75), c)
    temps = np.geomspace(0.2, 2.0, 32)
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, seed=3, tune=True, tune_sweeps=4)
    t = mc.tuning
    assert t["cached"] == 0 and t["n_candidates"] >= 3
    assert t["threads"] in (256, 512, 1024)
    assert 0 < t["ms_per_sweep"] <= t["default_ms"]
    assert t["colour_cap"] == 0 or t["colour_cap"] % 64 == 0 or t["colour_cap"] == lat.N
    # the tuned colouring is a proper colouring of the active couplings
    col = lat.colours()
    nb = lat.tables()
    for i in range(lat.N):
        for j in nb[i]:
            if j >= 0:
                assert col[i] != col[j]
    mc.set_spins(_spins(rng, len(temps), lat.N))
    mc.run(20, pt_parity=2)
    assert np.all(np.isfinite(mc.energies()))
    # second call: from the cache, same choice, nothing measured
    lat2 = H.Lattice.from_points(pts, 30.0, (1.15, 1.45, 1.75), c)
    mc2 = H.MonteCarlo(lat2, temps, precision=H.BATCHED_F32, seed=3, tune=True, tune_sweeps=4)
    assert mc2.tuning["cached"] == 1
    for k in ("colour_cap", "threads"):
        assert mc2.tuning[k] == t[k]


@pytest.mark.parametrize("cap,threads", [(32, 256), (64, 512), (0, 64)])
def test_ising_with_any_schedule_and_shape_is_bit_exact(cap, threads):
    _say("ising schedule", cap, threads)
    c = H.Couplings(J1=1.0, J2=0.4, JX=0.1, h=0.3)
    lat = H.Lattice.from_points(H.poisson_disk(20.0, 1.0, seed=4), 20.0, (1.15, 1.45, 1.75), c)
    lat.set_colour_cap(cap)
    m = cpu_ref.Model(lat.tables(), (lat.zn1, lat.zn2, lat.znx), c.as_params())
    temps = np.geomspace(0.5, 5.0, 8)
    R = len(temps)
    rng = np.random.default_rng(7)
    s = np.zeros((R, lat.N, 3))
    s[:, :, 0] = rng.choice([-1.0, 1.0], size=(R, lat.N))
    mc = H.MonteCarlo(lat, temps, model=H.ISING, precision=H.BATCHED_F32, seed=5)
    mc.set_threads(threads)
    mc.set_spins(s)
    tr = mc.run(12, pt_parity=2, trace=True)
    ref = s.copy()
    rtr, (cos, cnt, exc, boost) = m.mc_replay(1, temps, lat.colours(), ref, 12, seed=5, pt_parity=2)
    np.testing.assert_array_equal(tr, rtr)
    np.testing.assert_array_equal(mc.slot_map(), cos)
    np.testing.assert_array_equal(mc.get_spins(), ref[cos])
    if cap:
        assert np.bincount(lat.colours()).max() <= cap


def test_heisenberg_threads_do_not_change_the_chain_beyond_rounding():
    _say("heis threads")
    c = H.Couplings(J1=1.0, J2=0.6, h=0.2)
    lat = H.Lattice.generate("square", 16, c)
    temps = np.array([0.3, 0.7, 1.5, 3.0])
    rng = np.random.default_rng(5)
    s = _spins(rng, 4, lat.N)
    out = []
    for thr in (1024, 256):
        mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F64, n_or=2, n_hb=2, seed=7)
        mc.set_threads(thr)
        mc.set_spins(s)
        mc.run(3, pt_parity=-1)
        out.append(mc.get_spins())
    np.testing.assert_allclose(out[0], out[1], atol=1e-12)


def test_smd_tuner_and_shapes_keep_parity(tmp_path, monkeypatch):
    _say("smd tuner")
    monkeypatch.setenv("HSMD_TUNE_CACHE", str(tmp_path / "tune.txt"))
    c = H.Couplings(J1=1.0, J2=0.6)
    lats = [H.Lattice.generate("square", 16, c), H.Lattice.generate("hexagn", 6, c)]
    idx = np.array([0, 1, 0, 1, 1, 0], np.int32)
    rng = np.random.default_rng(9)
    samples = [_spins(rng, 1, lats[i].N)[0] for i in idx]
    T, dt = 256, 0.005
    sd = H.SpinDynamics(lats, idx, dt, T, n_omega=128, tune=True, tune_steps=16)
    t = sd.tuning
    assert t["cached"] == 0 and t["n_candidates"] >= 3
    assert t["form"] in (1, 2, 3) and t["gemm"] in (0, 1, 2, 3)   # 2 n_omega = 256: no 512 tile
    assert t["ms_per_step"] <= t["default_ms_per_step"] and t["gemm_ms"] <= t["default_gemm_ms"]
    sd2 = H.SpinDynamics(lats, idx, dt, T, n_omega=128, tune=True, tune_steps=16)
    assert sd2.tuning["cached"] == 1
    # every shape gives the same dynamics and DSSF up to FP32 rounding
    res = []
    for form, gemm in [(0, 0), (1, 1), (2, 2), (3, 1), (2, 0), (1, 3)]:
        e = H.SpinDynamics(lats, idx, dt, T, n_omega=128)
        e.set_shape(form, gemm)
        e.set_samples(samples)
        e.step(T - 1)
        res.append((np.concatenate(e.get_samples()), e.dssf(0)[0], e.dssf(1)[0]))
    for r in res[1:]:
        np.testing.assert_allclose(r[0], res[0][0], atol=1e-4)
        for k in (1, 2):
            np.testing.assert_allclose(r[k], res[0][k], rtol=1e-3, atol=1e-6 * res[0][k].max())


def test_smd_tuner_with_the_fft_form(tmp_path, monkeypatch):
    """n_steps = 1024 (a multiple of 512) and 128 bins: the FFT form is the
    default and a tuner candidate; the chosen shape and every transform form
    give the same DSSF up to FP32 rounding."""
    _say("smd tuner with the fft form")
    monkeypatch.setenv("HSMD_TUNE_CACHE", str(tmp_path / "tune.txt"))
    c = H.Couplings(J1=1.0, J2=0.6)
    lats = [H.Lattice.generate("square", 16, c), H.Lattice.generate("hexagn", 6, c)]
    idx = np.array([0, 1, 0, 1], np.int32)
    rng = np.random.default_rng(19)
    samples = [_spins(rng, 1, lats[i].N)[0] for i in idx]
    T, dt = 1024, 0.005
    sd = H.SpinDynamics(lats, idx, dt, T, n_omega=128, tune=True, tune_steps=16)
    t = sd.tuning
    assert t["form"] in (1, 2, 3) and t["gemm"] in (0, 1, 2, 3, 6)
    assert t["gemm_ms"] <= t["default_gemm_ms"]
    res = []
    for gemm in (6, 0, t["gemm"]):
        e = H.SpinDynamics(lats, idx, dt, T, n_omega=128)
        e.set_shape(t["form"], gemm)
        e.set_samples(samples)
        e.step(T - 1)
        res.append((e.dssf(0)[0], e.dssf(1)[0]))
    for r in res[1:]:
        for k in (0, 1):
            np.testing.assert_allclose(r[k], res[0][k], rtol=1e-3, atol=1e-6 * res[0][k].max())


def test_shape_errors():
    _say("shape errors")
    c = H.Couplings(J1=1.0, J2=0.6)
    lat = H.Lattice.generate("square", 8, c)
    e = H.SpinDynamics([lat], np.zeros(2, np.int32), 0.01, 64, n_omega=64)
    with pytest.raises(H.HSMDError):
        e.set_shape(2, 1)      # one slot class only
    with pytest.raises(H.HSMDError):
        e.set_shape(0, 4)      # 2 n_omega = 128: no 512-column tile
    with pytest.raises(H.HSMDError):
        e.set_shape(0, 5)
    with pytest.raises(H.HSMDError):
        e.set_shape(0, 6)      # n_steps = 64: no FFT form (needs a multiple of 512)
    with pytest.raises(H.HSMDError):
        e.set_shape(0, 7)
    e64 = H.SpinDynamics([lat], np.zeros(2, np.int32), 0.01, 64, n_omega=64, precision=H.BATCHED_F64)
    with pytest.raises(H.HSMDError):
        e64.set_shape(1, 0)    # the 128 x 128 tile is FP32 only
    e64b = H.SpinDynamics([lat], np.zeros(2, np.int32), 0.01, 512, n_omega=64, precision=H.BATCHED_F64)
    with pytest.raises(H.HSMDError):
        e64b.set_shape(0, 6)   # the FFT form is FP32 only
    mc = H.MonteCarlo(lat, [1.0], seed=1)
    with pytest.raises(H.HSMDError):
        mc.set_threads(100)
