"""GPU tests of the fast sweep kernels' wave-level class ordering.

With a neighbour-owner mask per (colour class, wave), a wave starts a class
once every wave owning one of its sites' neighbours has finished the class
before; under a workgroup barrier per class each site reads the same values.
So the two orderings must give bit-identical chains: spins, FP64 energies,
slot maps and exchange counters, through the sweeps, the heat-bath draws and
the replica exchange (the reference's sweep and exchange loop,
CF/mc_hm_code.cpp:2778-3130).  HSMD_MC_WAVE_SYNC = 1 / 0 forces either
ordering at engine creation."""
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
