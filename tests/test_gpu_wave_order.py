"""GPU tests of the wave-level ordering of the fast sweep kernels and the
LDS-resident RK4 kernel.

With a neighbour-owner mask per (colour class, wave), a wave starts a class
once every wave owning one of its sites' neighbours has finished the class
before; under a workgroup barrier per class each site reads the same values.
So the two orderings must give bit-identical chains: spins, FP64 energies,
slot maps and exchange counters, through the sweeps, the heat-bath draws and
the replica exchange (the reference's sweep and exchange loop,
CF/mc_hm_code.cpp:2778-3130), for the Heisenberg and the Ising kernels; and
the RK4 stages ordered per wave must give the trajectories, energies, qSS
and DSSF of a barrier per stage (CF/subrouts.cpp:11734-11902).
HSMD_MC_WAVE_SYNC / HSMD_SMD_WAVE_SYNC = 1 / 0 force either ordering at
engine creation."""
import math
import sys

import numpy as np
import pytest

import classical_mc_and_smd_for_2d_heisenberg_models_amd as H

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-wave]", *a, flush=True)
    sys.stdout.flush()


def _unit(rng, shape):
    v = rng.normal(size=shape)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _chain(monkeypatch, mode, lat, temps, spins, n_sweeps, threads=None, **kw):
    monkeypatch.setenv("HSMD_MC_WAVE_SYNC", mode)
    mc = H.MonteCarlo(lat, temps, **kw)
    if threads:
        mc.set_threads(threads)
    name = mc.sweep_kernel()
    mc.set_spins(spins)
    tr = mc.run(n_sweeps, pt_parity=2, trace=True)
    out = (name, mc.get_spins(), mc.energies(), tr, mc.slot_map(), mc.pt_counters())
    del mc
    return out


def _compare(monkeypatch, lat, temps, spins, n_sweeps, threads=None, **kw):
    a = _chain(monkeypatch, "1", lat, temps, spins, n_sweeps, threads, **kw)
    b = _chain(monkeypatch, "0", lat, temps, spins, n_sweeps, threads, **kw)
    _say(a[0], "|", b[0])
    assert a[0].endswith("wave-ordered"), a[0]
    assert not b[0].endswith("wave-ordered"), b[0]
    assert a[0].startswith("heis_sweep_fast"), a[0]
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])
    np.testing.assert_array_equal(a[3], b[3])
    np.testing.assert_array_equal(a[4], b[4])
    np.testing.assert_array_equal(a[5][0], b[5][0])
    np.testing.assert_array_equal(a[5][1], b[5][1])
    assert np.isfinite(a[2]).all()
    return a


def test_config3_wave_ordered_sweeps_equal_barrier_sweeps(monkeypatch):
    """Triangular L=96, 256 FP32 replicas (the statically placed float4
    kernel with the half-shell table energy), 5 OR + 1 HB, the (2, 8) boost
    after accepted swaps, 4 sweeps with exchange; the default ordering of this
    lattice is the wave-ordered one."""
    _say("config3 wave order")
    c = H.Couplings(J1=1.0, J2=0.12, JX=0.05, h=0.3, lambda1=0.9, lambda2=0.9)
    lat = H.Lattice.generate("triang", 96, c)
    temps = np.geomspace(0.02, 2.0, 256)
    s = _unit(np.random.default_rng(41), (256, lat.N, 3))
    a = _compare(monkeypatch, lat, temps, s, 4, precision=H.BATCHED_F32, n_or=5, n_hb=1, seed=9)
    assert "true,false,true,5" in a[0]
    monkeypatch.delenv("HSMD_MC_WAVE_SYNC")
    mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, n_or=5, n_hb=1, seed=9)
    assert mc.sweep_kernel().endswith("wave-ordered")
    np.testing.assert_allclose(np.linalg.norm(a[1], axis=-1), 1.0, atol=1e-5)


@pytest.mark.parametrize("prec", [H.BATCHED_F32, H.BATCHED_F64])
def test_config1_wave_ordered_sweeps_equal_barrier_sweeps(monkeypatch, prec):
    """Square L=24 J1-J2, 16 replicas, 1 HB + 5 OR: the float4 kernel placed
    at run time (FP32) and the component-array FP64 kernel; 6 sweeps."""
    _say("config1 wave order", prec)
    c = H.Couplings(J1=1.0, J2=0.6)
    lat = H.Lattice.generate("square", 24, c)
    temps = np.geomspace(0.05, 1.5, 16)
    s = _unit(np.random.default_rng(42), (16, lat.N, 3))
    _compare(monkeypatch, lat, temps, s, 6, precision=prec, n_or=5, n_hb=1, seed=4)


@pytest.mark.parametrize("threads", [512, 256])
def test_narrow_workgroups_wave_ordered_equal_barrier(monkeypatch, threads):
    """Triangular L=48 (2304 sites) with classes capped to the workgroup
    width: 512 and 256 threads, 8 and 4 waves per replica."""
    _say("narrow workgroups", threads)
    c = H.Couplings(J1=1.0, J2=0.12, JX=0.05, h=0.3)
    lat = H.Lattice.generate("triang", 48, c)
    lat.set_colour_cap(threads)
    temps = np.geomspace(0.1, 2.0, 32)
    s = _unit(np.random.default_rng(43), (32, lat.N, 3))
    _compare(monkeypatch, lat, temps, s, 3, threads=threads, precision=H.BATCHED_F32, n_or=2,
             n_hb=1, seed=5)


def test_poisson_graph_compact_labels_wave_ordered_equal_barrier(monkeypatch):
    """The ~12,000-site Poisson-disk graph with 5 % vacancies (compact
    labels, 15 classes of at most 1024 sites), ordering forced on: 128 FP32
    replicas, 2 sweeps."""
    _say("graph wave order")
    box = math.sqrt(12000 / 0.627)
    pts = H.poisson_disk(box, 1.0, seed=5)
    c = H.Couplings(J1=1.0, J2=0.4, JX=0.1)
    lat = H.Lattice.from_points(pts, box, (1.15, 1.45, 1.75), c)
    rng = np.random.default_rng(505)
    vac = np.zeros(lat.N, np.int32)
    vac[rng.choice(lat.N, int(round(0.05 * lat.N)), replace=False)] = 1
    lat.set_vacancies(vac)
    temps = np.geomspace(0.05, 1.5, 128)
    s = _unit(rng, (128, lat.N, 3))
    a = _compare(monkeypatch, lat, temps, s, 2, precision=H.BATCHED_F32, n_or=5, n_hb=1, seed=6)
    assert ",true> x" in a[0]


# ------------------------------------------------------------ SMD RK4

def _smd_chain(monkeypatch, mode, lats, idx, samples, T, form=None, **kw):
    monkeypatch.setenv("HSMD_SMD_WAVE_SYNC", mode)
    smd = H.SpinDynamics(lats, idx, 0.005, T, **kw)
    if form:
        smd.set_shape(form, 1)
    name = smd.step_kernel()
    smd.set_samples(samples)
    smd.step(T - 1)
    out = [name, smd.get_samples(), smd.energies()]
    if kw.get("dssf", True):
        for li in range(len(lats)):
            out.extend(smd.dssf(li))
    del smd
    return out


def _smd_compare(monkeypatch, lats, idx, samples, T, form=None, **kw):
    a = _smd_chain(monkeypatch, "1", lats, idx, samples, T, form, **kw)
    b = _smd_chain(monkeypatch, "0", lats, idx, samples, T, form, **kw)
    _say(a[0], "|", b[0])
    assert "wave-ordered on %d of %d lattices" % (len(lats), len(lats)) in a[0], a[0]
    assert "wave-ordered" not in b[0], b[0]
    assert len(a[1]) == len(b[1]) == len(idx)
    for x, y in zip(a[1], b[1]):   # samples of uneven size
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a[2:], b[2:]):
        np.testing.assert_array_equal(x, y)
    assert np.isfinite(a[2]).all()
    return a


@pytest.mark.parametrize("form", [2, 1])
def test_config4_wave_ordered_rk4_equals_barrier_rk4(monkeypatch, form):
    """The four configuration-4 lattices (square L=64, kagome L=32,
    hexagonal L=40, square L=48 with 230 vacancies), two FP32 samples each,
    four sites per thread: per-workgroup slot classes (smd_kernel_two) and
    the batch's largest class; 1024 RK4 steps (a launch boundary at 1000),
    trajectories, energies, qSS and the 64-bin DSSF bit-identical."""
    _say("config4 smd wave order, form", form)
    rng = np.random.default_rng(411)
    c = H.Couplings(J1=1.0, J2=0.6)
    lats = [H.Lattice.generate("square", 64, c), H.Lattice.generate("kagome", 32, c),
            H.Lattice.generate("hexagn", 40, c), H.Lattice.generate("square", 48, c)]
    vac = np.zeros(lats[3].N, np.int32)
    vac[rng.choice(lats[3].N, 230, replace=False)] = 1
    lats[3].set_vacancies(vac)
    idx = np.array([0, 1, 2, 3, 3, 2, 1, 0], np.int32)
    samples = [_unit(rng, (lats[i].N, 3)) for i in idx]
    a = _smd_compare(monkeypatch, lats, idx, samples, 1025, form=form, n_omega=64,
                     precision=H.BATCHED_F32)
    assert ("smd_kernel_two<float,4,8,10>" in a[0]) == (form == 2), a[0]


@pytest.mark.parametrize("prec", [H.BATCHED_F32, H.BATCHED_F64])
def test_small_lattices_wave_ordered_rk4_equals_barrier_rk4(monkeypatch, prec):
    """Square L=24 (one site per thread: the configuration-2 lattice) and
    triangular L=36 (1296 sites, two per thread; FP64 with both in one
    batch at two sites per thread), 300 RK4 steps with the 32-bin DSSF."""
    _say("small lattices smd wave order", prec)
    rng = np.random.default_rng(412)
    lats = [H.Lattice.generate("square", 24, H.Couplings(J1=1.0, J2=0.6)),
            H.Lattice.generate("triang", 36, H.Couplings(J1=1.0, J2=0.12, JX=0.05))]
    idx = np.array([0, 1, 1, 0], np.int32)
    samples = [_unit(rng, (lats[i].N, 3)) for i in idx]
    _smd_compare(monkeypatch, lats, idx, samples, 301, n_omega=32, precision=prec)


# ------------------------------------------------------------ Ising

def test_config3_ising_fast_kernel_wave_ordered_equals_barrier_and_oracle(monkeypatch):
    """The Ising variant of configuration 3 (int8 spins, h = 0.3) on the fast
    Ising kernel: the energy from the half-shell table (verified equal to the
    bond list as a multiset), classes ordered per wave or by a barrier; both
    bit-identical, and the energy trace, slot map and spins bit-exact
    against the oracle's colour replay over four sweeps with exchange."""
    import cpu_ref
    _say("config3 ising fast")
    c = H.Couplings(J1=1.0, J2=0.12, JX=0.05, h=0.3, lambda1=0.9, lambda2=0.9)
    lat = H.Lattice.generate("triang", 96, c)
    temps = np.geomspace(0.02, 2.0, 256)
    rng = np.random.default_rng(413)
    s = np.zeros((256, lat.N, 3))
    s[:, :, 0] = rng.choice([-1.0, 1.0], size=(256, lat.N))
    a = _compare_ising(monkeypatch, lat, temps, s, 4, model=H.ISING, precision=H.BATCHED_F32, seed=29)
    assert a[0].startswith("ising_sweep_fast<20> x 1024, table energy"), a[0]
    m = cpu_ref.Model(lat.tables(), (lat.zn1, lat.zn2, lat.znx), c.as_params())
    ref = s.copy()
    rtr, (cos, cnt, exc, _) = m.mc_replay(1, temps, lat.colours(), ref, 4, seed=29, pt_parity=2)
    np.testing.assert_array_equal(a[3], rtr)
    np.testing.assert_array_equal(a[4], cos)
    np.testing.assert_array_equal(a[1], ref[cos])
    np.testing.assert_array_equal(a[5][1], exc)


def _compare_ising(monkeypatch, lat, temps, spins, n_sweeps, **kw):
    a = _chain(monkeypatch, "1", lat, temps, spins, n_sweeps, None, **kw)
    b = _chain(monkeypatch, "0", lat, temps, spins, n_sweeps, None, **kw)
    _say(a[0], "|", b[0])
    assert a[0].endswith("wave-ordered") and not b[0].endswith("wave-ordered")
    for x, y in zip(a[1:5], b[1:5]):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a[5][0], b[5][0])
    np.testing.assert_array_equal(a[5][1], b[5][1])
    return a


@pytest.mark.parametrize("geom,L,kw", [("square", 4, dict(J1=1.0, J2=0.6, JX=0.3)),
                                       ("square", 24, dict(J1=1.0, J2=0.6)),
                                       ("kagome", 6, dict(J1=1.0, J2=0.4, h=0.2)),
                                       ("hexagn", 6, dict(J1=1.0, J2=0.3, JX=0.2))])
def test_ising_fast_kernel_small_lattices_bit_exact_against_oracle(geom, L, kw):
    """The fast Ising kernel on small lattices, among them square L=4 with a
    third shell (whose doubled PBC neighbours leave the half-shell table
    unequal to the bond list, so the bond-list energy stays): 16 replicas,
    six sweeps with exchange, bit-exact against the oracle's colour replay."""
    import cpu_ref
    c = H.Couplings(**kw)
    lat = H.Lattice.generate(geom, L, c)
    temps = np.geomspace(0.3, 4.0, 16)
    rng = np.random.default_rng(414)
    s = np.zeros((16, lat.N, 3))
    s[:, :, 0] = rng.choice([-1.0, 1.0], size=(16, lat.N))
    mc = H.MonteCarlo(lat, temps, model=H.ISING, precision=H.BATCHED_F32, seed=31)
    name = mc.sweep_kernel()
    _say(geom, L, name)
    assert name.startswith("ising_sweep_fast"), name
    mc.set_spins(s)
    tr = mc.run(6, pt_parity=2, trace=True)
    m = cpu_ref.Model(lat.tables(), (lat.zn1, lat.zn2, lat.znx), c.as_params())
    ref = s.copy()
    rtr, (cos, cnt, exc, _) = m.mc_replay(1, temps, lat.colours(), ref, 6, seed=31, pt_parity=2)
    np.testing.assert_array_equal(tr, rtr)
    np.testing.assert_array_equal(mc.slot_map(), cos)
    np.testing.assert_array_equal(mc.get_spins(), ref[cos])
