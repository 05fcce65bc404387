"""The replayed-graph sweep loop (hsmd_mc_run with n_sweeps > 1 and no energy
trace) against the launch-by-launch loop (taken when an energy trace is
asked for): same kernels, same counters, so spins, slot maps, exchange
counters and measured moments must agree bit for bit.  The graph engine is
replayed in several calls with eager work (energy reads, a one-sweep call)
between them, and re-captured when the parity mode or measurement changes."""
import sys

import numpy as np
import pytest

import classical_mc_and_smd_for_2d_heisenberg_models_amd as H

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-graph]", *a, flush=True)
    sys.stdout.flush()


def _start(lat, model, R, rng):
    if model == H.ISING:
        s = np.zeros((R, lat.N, 3))
        s[:, :, 0] = rng.choice([-1.0, 1.0], size=(R, lat.N))
        return s
    v = rng.normal(size=(R, lat.N, 3))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


@pytest.mark.parametrize("model,prec", [(H.HEISENBERG, H.BATCHED_F32), (H.HEISENBERG, H.BATCHED_F64),
                                        (H.ISING, H.BATCHED_F32)])
@pytest.mark.parametrize("pt,measure", [(2, False), (-1, False), (0, True), (2, True)])
def test_graph_sweep_loop_matches_launch_by_launch(model, prec, pt, measure):
    _say("graph vs eager", model, prec, pt, measure)
    c = H.Couplings(J1=1.0, J2=0.12, JX=0.05, h=0.2)
    lat = H.Lattice.generate("triang", 12, c)
    temps = np.geomspace(0.3, 3.0, 8)
    rng = np.random.default_rng(11)
    s0 = _start(lat, model, len(temps), rng)

    def engine():
        mc = H.MonteCarlo(lat, temps, model=model, precision=prec, n_or=3, n_hb=2, seed=41)
        mc.set_spins(s0)
        return mc

    eager = engine()
    eager.run(30, pt_parity=pt, measure=measure, trace=True)
    graph = engine()
    graph.run(10, pt_parity=pt, measure=measure)
    graph.energies()
    graph.run(10, pt_parity=pt, measure=measure)
    graph.run(1, pt_parity=pt, measure=measure)
    graph.slot_map()
    graph.run(9, pt_parity=pt, measure=measure)
    np.testing.assert_array_equal(graph.get_spins(), eager.get_spins())
    np.testing.assert_array_equal(graph.slot_map(), eager.slot_map())
    np.testing.assert_array_equal(graph.energies(), eager.energies())
    for a, b in zip(graph.pt_counters(), eager.pt_counters()):
        np.testing.assert_array_equal(a, b)
    if measure:
        mg, ng = graph.moments()
        me, ne = eager.moments()
        assert ng == ne == 30
        np.testing.assert_array_equal(mg, me)
    if pt >= 0:
        assert eager.pt_counters()[1].sum() > 0


def test_graph_recaptures_when_the_mode_changes():
    _say("graph re-capture")
    c = H.Couplings(J1=1.0, J2=0.5)
    lat = H.Lattice.generate("square", 16, c)
    temps = np.geomspace(0.2, 2.0, 6)
    rng = np.random.default_rng(3)
    s0 = _start(lat, H.HEISENBERG, len(temps), rng)
    plan = [(5, 2, False), (4, -1, False), (6, 1, True), (3, 2, True), (5, 2, False)]
    out = []
    for trace in (False, True):
        mc = H.MonteCarlo(lat, temps, precision=H.BATCHED_F32, seed=8)
        mc.set_spins(s0)
        for n, pt, meas in plan:
            mc.run(n, pt_parity=pt, measure=meas, trace=trace)
        out.append((mc.get_spins(), mc.slot_map(), mc.moments()[0]))
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
