"""GPU parity tests of the SMD engine (RK4 + energy correction + DSSF)."""
import sys

import numpy as np
import pytest

import classical_mc_and_smd_for_2d_heisenberg_models_amd as H
import cpu_ref

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-smd]", *a, flush=True)
    sys.stdout.flush()


def _spins(rng, N):
    v = rng.normal(size=(N, 3))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _model(lat, c):
    return cpu_ref.Model(lat.tables(), (lat.zn1, lat.zn2, lat.znx), c.as_params())


@pytest.mark.parametrize("prec,tol", [(H.BATCHED_F64, 1e-10), (H.PLAIN_F64, 1e-10),
                                      (H.BATCHED_F32, 2e-5)])
def test_trajectories_match_oracle_over_100_steps(prec, tol):
    _say("trajectory", prec)
    c = H.Couplings(J1=1.0, J2=0.6)
    lat = H.Lattice.generate("square", 8, c)
    m = _model(lat, c)
    rng = np.random.default_rng(4)
    samples = [_spins(rng, lat.N) for _ in range(3)]
    T = 101
    smd = H.SpinDynamics([lat], [0, 0, 0], 0.01, T, n_omega=32, precision=prec, dssf=False)
    smd.set_samples(samples)
    smd.step(T - 1)
    out = smd.get_samples()
    E = smd.energies()
    for k, s0 in enumerate(samples):
        s_ref, tr, q, e = m.smd_run(s0, 0.01, T, gauss_seidel=True)
        rel = np.abs(out[k] - s_ref).max()
        assert rel < tol, (k, rel)
        assert E[k] == pytest.approx(e[-1], rel=max(tol, 1e-9))


def test_mixed_lattice_batch_matches_oracle():
    _say("mixed batch")
    cs = [H.Couplings(J1=1.0, J2=0.6), H.Couplings(J1=1.0, J2=0.1, lambda1=0.9, lambda2=0.9),
          H.Couplings(J1=1.0, J2=0.2, JX=0.1, h=0.3)]
    lats = [H.Lattice.generate("square", 6, cs[0]), H.Lattice.generate("kagome", 4, cs[1]),
            H.Lattice.generate("hexagn", 4, cs[2])]
    vac = np.zeros(lats[0].N, np.int32)
    vac[[3, 17, 20]] = 1
    lats[0].set_vacancies(vac)
    idx = [0, 1, 2, 1, 0]
    rng = np.random.default_rng(6)
    samples = [_spins(rng, lats[i].N) for i in idx]
    T = 60
    smd = H.SpinDynamics(lats, idx, 0.01, T, n_omega=32, precision=H.BATCHED_F64, dssf=False)
    smd.set_samples(samples)
    smd.step(T - 1)
    out = smd.get_samples()
    for k, i in enumerate(idx):
        nb = lats[i].tables().copy()
        if i == 0:
            # vacancies: every entry touching a vacant site is a hole
            for v in np.nonzero(vac)[0]:
                nb[v, :] = -1
                nb[nb == v] = -1
        m = cpu_ref.Model(nb, (lats[i].zn1, lats[i].zn2, lats[i].znx), cs[i].as_params())
        s_ref = m.smd_run(samples[k], 0.01, T, gauss_seidel=False)[0]
        assert np.abs(out[k] - s_ref).max() < 1e-10, k


@pytest.mark.parametrize("geom,L", [("square", 4), ("kagome", 4), ("hexagn", 3)])
def test_dssf_fp64_matches_oracle(geom, L):
    _say("dssf fp64", geom, L)
    c = H.Couplings(J1=1.0, J2=0.3)
    lat = H.Lattice.generate(geom, L, c)
    m = _model(lat, c)
    rng = np.random.default_rng(7)
    S, T, no = 2, 64, 32
    samples = [_spins(rng, lat.N) for _ in range(S)]
    w = cpu_ref.window("hann1", T)
    smd = H.SpinDynamics([lat], [0] * S, 0.02, T, n_omega=no, precision=H.BATCHED_F64,
                         window=w)
    smd.set_samples(samples)
    smd.step(T - 1)
    d, q = smd.dssf(0)
    ref = np.zeros_like(d)
    qref = np.zeros(T)
    for s0 in samples:
        _, tr, qq, _ = m.smd_run(s0, 0.02, T, gauss_seidel=False)
        ref += cpu_ref.dssf(tr, lat.grid_map(), lat.Gsz, w, no)
        qref += qq
    np.testing.assert_allclose(d, ref, rtol=1e-8, atol=1e-10 * ref.max())
    np.testing.assert_allclose(q, qref, rtol=1e-10)


@pytest.mark.parametrize("no", [64, 128])
def test_dssf_fp32_within_1e3_of_fp64(no):
    # no = 64 runs the 64-wide MFMA tiles, no = 128 the 256-wide gfx950 tiles
    _say("dssf fp32 vs fp64", no)
    c = H.Couplings(J1=1.0, J2=0.6)
    lat = H.Lattice.generate("square", 8, c)
    rng = np.random.default_rng(8)
    S, T = 4, 300
    samples = [_spins(rng, lat.N) for _ in range(S)]
    out = {}
    for prec in (H.BATCHED_F64, H.BATCHED_F32):
        smd = H.SpinDynamics([lat], [0] * S, 0.02, T, n_omega=no, precision=prec)
        smd.set_samples(samples)
        smd.step(T - 1)
        out[prec] = smd.dssf(0)[0]
    a, b = out[H.BATCHED_F64], out[H.BATCHED_F32]
    err = np.abs(a - b).max() / a.max()
    _say("fp32 dssf max err / peak", err)
    assert err < 1e-3
