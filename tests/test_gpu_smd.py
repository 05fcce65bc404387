"""GPU parity tests of the SMD engine (RK4 + energy correction + DSSF)."""
import math
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


@pytest.mark.parametrize("prec,tol", [(H.BATCHED_F64, 1e-10), (H.BATCHED_F32, 5e-5)])
def test_mixed_lattice_batch_matches_oracle(prec, tol):
    # FP32 runs each workgroup with its own lattice's slot class (8 or 12)
    _say("mixed batch", prec)
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
    smd = H.SpinDynamics(lats, idx, 0.01, T, n_omega=32, precision=prec, dssf=False)
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
        assert np.abs(out[k] - s_ref).max() < tol, k


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


@pytest.mark.parametrize("max_chunk", [2048, None])
def test_dssf_fp32_across_trajectory_chunks_within_1e3_of_fp64(max_chunk, monkeypatch):
    # 4500 time points: with the FP32 chunk capped at 2048 points, two full
    # chunks (the first stores the contraction, the second adds to it) and a
    # 404-point tail (not a multiple of the contraction's 32-point step);
    # uncapped, one 4500-point chunk; against the FP64 form (256-point
    # chunks); dt = 0.002 keeps the two precisions' trajectories close over
    # the whole run
    if max_chunk:
        monkeypatch.setenv("HSMD_SMD_MAX_CHUNK", str(max_chunk))
    _say("dssf fp32 vs fp64 across chunks", 4500, "max chunk", max_chunk)
    c = H.Couplings(J1=1.0, J2=0.6)
    lat = H.Lattice.generate("square", 8, c)
    rng = np.random.default_rng(18)
    S, T = 3, 4500
    samples = [_spins(rng, lat.N) for _ in range(S)]
    out = {}
    for prec in (H.BATCHED_F64, H.BATCHED_F32):
        smd = H.SpinDynamics([lat], [0] * S, 0.002, T, n_omega=128, precision=prec)
        smd.set_samples(samples)
        smd.step(2047)
        smd.step(T - 1 - 2047)
        out[prec] = smd.dssf(0)
    (a, qa), (b, qb) = out[H.BATCHED_F64], out[H.BATCHED_F32]
    err = np.abs(a - b).max() / a.max()
    _say("fp32 dssf max err / peak", err)
    assert err < 1e-3
    # qSS is formed every step (no chunking); its FP32 drift over 4500 steps
    # is held to the same 1e-3 of its peak
    np.testing.assert_allclose(qb, qa, rtol=0, atol=1e-3 * qa.max())


def _qc_fixture(n):
    import os
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "qc_with_PBC_%d" % n)
    nb = np.fromfile(os.path.join(d, "nborsList_AppxtSz(%d).bin" % n), dtype="<i4").reshape(n, 19)
    rv = np.fromfile(os.path.join(d, "sitesList_AppxtSz(%d).bin" % n), dtype="<f8").reshape(n, 2)
    return nb, rv


@pytest.mark.parametrize("prec,tol", [(H.BATCHED_F64, 1e-8), (H.BATCHED_F32, 2e-4)])
def test_dssf_direct_sums_on_tabulated_graph_match_oracle(prec, tol):
    # tabulated graphs take get_qct_SqData's direct sums on the L x L q grid
    _say("dssf direct tabulated", prec)
    nb, rv = _qc_fixture(41)
    c = H.Couplings(J1=1.0, J2=0.4, JX=0.1)
    lat = H.Lattice.from_tables(nb, 8, 3, 8, rv, c)
    m = cpu_ref.Model(nb, (8, 3, 8), c.as_params())
    rng = np.random.default_rng(12)
    S, T, no = 2, 64, 32
    samples = [_spins(rng, lat.N) for _ in range(S)]
    w = cpu_ref.window("hann1", T)
    smd = H.SpinDynamics([lat], [0] * S, 0.02, T, n_omega=no, precision=prec, window=w)
    smd.set_samples(samples)
    smd.step(T - 1)
    d, _ = smd.dssf(0)
    Lq = lat.L
    dq = 4 * 2 * math.pi / lat.Gsz
    qv = np.zeros((lat.Nq, 2))
    for j in range(Lq):
        for i in range(Lq):
            nx = i if i < Lq // 2 else i - Lq
            ny = j if j < Lq // 2 else j - Lq
            qv[j * Lq + i] = (dq * nx, dq * ny)
    ref = np.zeros_like(d)
    for s0 in samples:
        _, tr, _, _ = m.smd_run(s0, 0.02, T, gauss_seidel=False)
        ref += cpu_ref.dssf_direct(tr, rv, qv, w, no)
    np.testing.assert_allclose(d, ref, rtol=tol, atol=tol * ref.max())
    # a q-path through the same lattice
    path = np.stack([np.linspace(0, 3.0, 17), np.linspace(0, -1.5, 17)], 1)
    dp = smd.dssf_q(path)
    refp = np.zeros_like(dp)
    for s0 in samples:
        _, tr, _, _ = m.smd_run(s0, 0.02, T, gauss_seidel=False)
        refp += cpu_ref.dssf_direct(tr, rv, path, w, no)
    np.testing.assert_allclose(dp, refp, rtol=tol, atol=tol * refp.max())


def test_dssf_q_on_a_periodic_lattice_with_vacancies_matches_oracle():
    _say("dssf_q square with vacancies")
    c = H.Couplings(J1=1.0, J2=0.5)
    lat = H.Lattice.generate("square", 6, c)
    rng = np.random.default_rng(13)
    vac = (rng.random(lat.N) < 0.15).astype(np.int32)
    lat.set_vacancies(vac)
    S, T, no = 2, 48, 32
    samples = [_spins(rng, lat.N) for _ in range(S)]
    w = np.ones(T)
    smd = H.SpinDynamics([lat], [0] * S, 0.02, T, n_omega=no, precision=H.BATCHED_F64)
    smd.set_samples(samples)
    smd.step(T - 1)
    path = np.stack([np.linspace(0, math.pi, 9), np.linspace(0, math.pi, 9)], 1)
    got = smd.dssf_q(path)
    rv = lat.positions()
    ref = np.zeros_like(got)
    # vacancies: every table entry touching a vacant site is a hole
    nb = lat.tables().copy()
    for v in np.nonzero(vac)[0]:
        nb[v, :] = -1
        nb[nb == v] = -1
    m = cpu_ref.Model(nb, (lat.zn1, lat.zn2, lat.znx), c.as_params())
    for k in range(S):
        _, tr, _, _ = m.smd_run(samples[k], 0.02, T, gauss_seidel=False)
        ref += cpu_ref.dssf_direct(tr, rv, path, w, no, mask=vac)
    np.testing.assert_allclose(got, ref, rtol=1e-8, atol=1e-10 * ref.max())


@pytest.mark.parametrize("prec,tol", [(H.BATCHED_F64, 1e-10), (H.BATCHED_F32, 2e-5)])
def test_wide_form_beyond_lds_matches_oracle(prec, tol):
    # N = 5184 exceeds the LDS-resident form: every array lives in global memory
    _say("wide form", prec)
    c = H.Couplings(J1=1.0, J2=0.6)
    lat = H.Lattice.generate("square", 72, c)
    m = _model(lat, c)
    rng = np.random.default_rng(14)
    samples = [_spins(rng, lat.N) for _ in range(2)]
    T = 41
    smd = H.SpinDynamics([lat], [0, 0], 0.01, T, n_omega=32, precision=prec, dssf=False)
    smd.set_samples(samples)
    smd.step(T - 1)
    out = smd.get_samples()
    E = smd.energies()
    for k in range(2):
        s_ref, _, _, e = m.smd_run(samples[k], 0.01, T, gauss_seidel=False, traj=False)
        assert np.abs(out[k] - s_ref).max() < tol, k
        assert E[k] == pytest.approx(e[-1], rel=max(tol, 1e-9))


@pytest.mark.parametrize("prec,tol", [(H.BATCHED_F64, 1e-10), (H.BATCHED_F32, 2e-5)])
def test_sparse_graph_wide_form_compact_labels_match_oracle(prec, tol):
    # a Poisson-disk graph beyond the LDS-resident form (N > 4096): the
    # global-memory form reads its sparse table as compact labels
    _say("sparse graph wide form", prec)
    box = 90.0
    pts = H.poisson_disk(box, 1.0, seed=11)
    c = H.Couplings(J1=1.0, J2=0.4, JX=0.1, h=0.15, lambda1=0.9, lambda2=1.1)
    lat = H.Lattice.from_points(pts, box, (1.15, 1.45, 1.75), c)
    assert lat.N > 4096
    m = _model(lat, c)
    rng = np.random.default_rng(16)
    samples = [_spins(rng, lat.N) for _ in range(2)]
    T = 31
    smd = H.SpinDynamics([lat], [0, 0], 0.01, T, n_omega=32, precision=prec, dssf=False)
    smd.set_samples(samples)
    smd.step(T - 1)
    out = smd.get_samples()
    E = smd.energies()
    for k in range(2):
        s_ref, _, _, e = m.smd_run(samples[k], 0.01, T, gauss_seidel=False, traj=False)
        assert np.abs(out[k] - s_ref).max() < tol, k
        assert E[k] == pytest.approx(e[-1], rel=max(tol, 1e-9))


def test_poisson_disk_graph_with_vacancies_matches_oracle():
    # an irregular tabulated graph (distance shells, 5% vacancies)
    _say("poisson graph smd")
    box = 40.0
    pts = H.poisson_disk(box, 1.0, seed=5)
    c = H.Couplings(J1=1.0, J2=0.4, JX=0.1)
    lat = H.Lattice.from_points(pts, box, (1.15, 1.45, 1.75), c)
    rng = np.random.default_rng(15)
    vac = (rng.random(lat.N) < 0.05).astype(np.int32)
    lat.set_vacancies(vac)
    nb = lat.tables().copy()
    for v in np.nonzero(vac)[0]:
        nb[v, :] = -1
        nb[nb == v] = -1
    m = cpu_ref.Model(nb, (lat.zn1, lat.zn2, lat.znx), c.as_params())
    samples = [_spins(rng, lat.N) for _ in range(3)]
    T = 51
    smd = H.SpinDynamics([lat], [0, 0, 0], 0.01, T, n_omega=32, precision=H.BATCHED_F64, dssf=False)
    smd.set_samples(samples)
    smd.step(T - 1)
    out = smd.get_samples()
    for k in range(3):
        s_ref = m.smd_run(samples[k], 0.01, T, gauss_seidel=False, traj=False)[0]
        assert np.abs(out[k] - s_ref).max() < 1e-10, k


@pytest.mark.parametrize("case", ["square_f32", "graph_f64", "ising_to_f32"])
def test_mc_to_smd_handoff_on_device_equals_host_round_trip(case):
    # the device hand-off (MC colour labels -> SMD labels, with precision
    # conversion) starts the dynamics exactly where set_samples with the MC
    # engine's spins does: same samples, same trajectory, same DSSF
    _say("handoff", case)
    c = H.Couplings(J1=1.0, J2=0.5)
    if case == "graph_f64":
        pts = H.poisson_disk(16.0, 1.0, seed=2)
        lat = H.Lattice.from_points(pts, 16.0, (1.15, 1.45, 1.75), H.Couplings(J1=1.0, J2=0.4, JX=0.1))
        mprec, sprec, model = H.BATCHED_F64, H.BATCHED_F64, H.HEISENBERG
    elif case == "square_f32":
        lat = H.Lattice.generate("square", 16, c)
        mprec, sprec, model = H.BATCHED_F32, H.BATCHED_F32, H.HEISENBERG
    else:
        lat = H.Lattice.generate("square", 16, c)
        mprec, sprec, model = H.BATCHED_F32, H.BATCHED_F32, H.ISING
    R = 4
    mc = H.MonteCarlo(lat, np.geomspace(0.3, 2.0, R), model=model, precision=mprec, seed=6)
    rng = np.random.default_rng(4)
    if model == H.ISING:
        s = np.zeros((R, lat.N, 3))
        s[..., 0] = np.where(rng.random((R, lat.N)) < 0.5, -1.0, 1.0)
    else:
        v = rng.normal(size=(R, lat.N, 3))
        s = v / np.linalg.norm(v, axis=-1, keepdims=True)
    mc.set_spins(s)
    mc.run(12, pt_parity=2)
    slots = np.array([3, 0, 3, 1, 2], np.int32)
    T = 64
    a = H.SpinDynamics([lat], np.zeros(5, np.int32), 0.01, T, n_omega=32, precision=sprec)
    b = H.SpinDynamics([lat], np.zeros(5, np.int32), 0.01, T, n_omega=32, precision=sprec)
    a.set_samples_from_mc(mc, slots)
    sp = mc.get_spins()
    b.set_samples([sp[k] for k in slots])
    for x, y in zip(a.get_samples(), b.get_samples()):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a.energies(), b.energies())
    a.step(T - 1)
    b.step(T - 1)
    for x, y in zip(a.get_samples(), b.get_samples()):
        np.testing.assert_array_equal(x, y)
    if case != "graph_f64":
        np.testing.assert_array_equal(a.dssf(0)[0], b.dssf(0)[0])
    # a sample on another lattice is refused
    other = H.Lattice.generate("square", 8, c)
    e = H.SpinDynamics([lat, other], np.array([0, 1], np.int32), 0.01, T, n_omega=32, precision=sprec)
    with pytest.raises(H.HSMDError):
        e.set_samples_from_mc(mc, np.array([0, 1], np.int32))
