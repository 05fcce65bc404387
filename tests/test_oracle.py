"""Pins the CPU oracle (oracle/cpu_ref.cpp) against known answers, closed
forms, textbook transforms and library routines, so that the GPU parity tests
compare against something that is itself checked.  CPU only."""
import math

import numpy as np
import pytest
from scipy.signal import windows as sw

import cpu_ref

PARAMS0 = np.array([1.0, 0, 0, 0, 1.0, 1.0, 0.0, 0, 0, 0, 0, 0, 0, 0])


def params(J1=1.0, J2=0.0, JX=0.0, J2d=0.0, l1=1.0, l2=1.0, h=0.0, imp=0.0,
           J1ic=0.0, J2ic=0.0, JXic=0.0, J1ii=0.0, J2ii=0.0, JXii=0.0):
    return np.array([J1, J2, JX, J2d, l1, l2, h, imp, J1ic, J2ic, JXic, J1ii, J2ii, JXii])


# ------------------------------------------------------------------ RNG

def test_philox_random123_kat():
    # Known-answer vectors of Philox4x32-10 published with Random123.
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, exp in kat:
        seed = key[0] | (key[1] << 32)
        out = cpu_ref.philox(seed, ctr)
        assert tuple(int(v) for v in out) == exp


def test_det_exp_matches_libm():
    rng = np.random.default_rng(3)
    xs = np.concatenate([-rng.random(20000) * 700, -rng.random(2000) * 1e-6, [0.0, -0.0, -1e-300]])
    for x in xs:
        y = cpu_ref.det_exp(x)
        e = math.exp(x)
        assert abs(y - e) <= 2 * math.ulp(e), (x, y, e)
    assert cpu_ref.det_exp(1.5) == 1.0
    assert cpu_ref.det_exp(-800.0) == 0.0


# -------------------------------------------------------------- lattices

@pytest.mark.parametrize("geom,L,zn", [("square", 6, (4, 4, 4)), ("triang", 6, (6, 6, 6)),
                                       ("hexagn", 6, (3, 6, 3)), ("kagome", 6, (4, 4, 6)),
                                       ("lieb", 6, (4, 4, 4))])
def test_oracle_lattice_shells_are_symmetric_and_at_right_distance(geom, L, zn):
    nb, rv, gm, z = cpu_ref.lattice(geom, L)
    assert z == zn
    N = nb.shape[0]
    # every site appears exactly once in the DFT grid map
    assert sorted(gm[gm >= 0].tolist()) == list(range(N))
    # shells are symmetric (j in shell(i) <=> i in shell(j))
    off = [0, zn[0], zn[0] + zn[1], sum(zn)]
    for s in range(3):
        for i in range(N):
            for j in nb[i, off[s]:off[s + 1]]:
                assert i in nb[j, off[s]:off[s + 1]]
    # the nearest shell sits at unit distance (minimum image)
    span = rv.max(axis=0) - rv.min(axis=0)
    assert np.all(span > 0)


def test_square_neighbour_order_matches_angle_sort():
    # square lattice, site (i, j) = i*L + j at (i, j); shell 1 sorted by
    # descending angle: up (pi/2), right (0)... atan2 ranges (-pi, pi]
    L = 6
    nb, rv, gm, zn = cpu_ref.lattice("square", L)
    i, j = 2, 3
    s = i * L + j
    up = i * L + (j + 1) % L        # +y, angle pi/2
    right = ((i + 1) % L) * L + j   # +x, angle 0
    down = i * L + (j - 1) % L      # -y, angle -pi/2
    left = ((i - 1) % L) * L + j    # -x, angle pi
    assert nb[s, :4].tolist() == [left, up, right, down]


# ------------------------------------------------- fields and energies

def _random_spins(rng, R, N):
    v = rng.normal(size=(R, N, 3))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def test_energy_closed_forms_square():
    L = 8
    nb, rv, gm, zn = cpu_ref.lattice("square", L)
    J1, J2 = 1.0, 0.6
    m = cpu_ref.Model(nb, zn, params(J1=J1, J2=J2))
    N = L * L
    ii, jj = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    neel = np.zeros((N, 3)); neel[:, 2] = np.cos(np.pi * (ii + jj)).ravel()
    stripe = np.zeros((N, 3)); stripe[:, 2] = np.cos(np.pi * jj).ravel()
    ferro = np.zeros((N, 3)); ferro[:, 0] = 1.0
    E = m.energies(np.stack([neel, stripe, ferro]))
    assert E[0] / N == pytest.approx(-2 * J1 + 2 * J2, abs=1e-13)
    assert E[1] / N == pytest.approx(-2 * J2, abs=1e-13)
    assert E[2] / N == pytest.approx((4 * J1 + 4 * J2) / 2, abs=1e-13)


def test_energy_closed_forms_triangular_120():
    L = 6
    nb, rv, gm, zn = cpu_ref.lattice("triang", L)
    m = cpu_ref.Model(nb, zn, params(J1=1.0))
    ii, jj = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    th = (4 * np.pi / 3) * (ii + 0.5 * jj)
    s = np.zeros((L * L, 3)); s[:, 0] = np.cos(th).ravel(); s[:, 1] = np.sin(th).ravel()
    assert m.energies(s)[0] / (L * L) == pytest.approx(-1.5, abs=1e-12)


def test_energy_ferromagnet_with_field_all_shells():
    for geom, L in [("triang", 6), ("kagome", 6), ("hexagn", 6)]:
        nb, rv, gm, zn = cpu_ref.lattice(geom, L)
        J1, J2, JX, h = 1.0, 0.12, 0.05, 0.3
        m = cpu_ref.Model(nb, zn, params(J1=J1, J2=J2, JX=JX, h=h))
        N = nb.shape[0]
        s = np.zeros((N, 3)); s[:, 2] = 1.0
        e = m.energies(s)[0] / N
        assert e == pytest.approx((zn[0] * J1 + zn[1] * J2 + zn[2] * JX) / 2 - h, abs=1e-12)


@pytest.mark.parametrize("geom,L,kw", [
    ("square", 8, dict(J1=1.0, J2=0.6)),
    ("triang", 6, dict(J1=1.0, J2=0.12, JX=0.05, h=0.3, l1=0.9, l2=0.9)),
    ("kagome", 4, dict(J1=1.0, J2=-0.2, JX=0.1, l1=1.2, l2=0.8, h=-0.4)),
    ("hexagn", 6, dict(J1=0.7, J2=0.3, JX=0.2)),
])
def test_bond_energy_equals_half_field_sum(geom, L, kw):
    # E = sum_bonds J S.S - h sum Sz  ==  1/2 sum_i S_i.h_i - 1/2 h sum Sz
    nb, rv, gm, zn = cpu_ref.lattice(geom, L)
    m = cpu_ref.Model(nb, zn, params(**kw))
    rng = np.random.default_rng(1)
    s = _random_spins(rng, 3, nb.shape[0])
    h = m.local_fields(s)
    E = m.energies(s)
    hz = kw.get("h", 0.0)
    Eh = 0.5 * np.einsum("rnd,rnd->r", s, h) - 0.5 * hz * s[:, :, 2].sum(axis=1)
    np.testing.assert_allclose(E, Eh, rtol=1e-12, atol=1e-11)


def test_local_field_is_energy_gradient():
    # h_i = dE/dS_i (finite differences), the defining property of the field
    nb, rv, gm, zn = cpu_ref.lattice("triang", 6)
    kw = dict(J1=1.0, J2=0.12, JX=0.05, h=0.3, l1=0.9, l2=0.7)
    m = cpu_ref.Model(nb, zn, params(**kw))
    rng = np.random.default_rng(2)
    s = _random_spins(rng, 1, nb.shape[0])[0]
    h = m.local_fields(s)[0]
    eps = 1e-6
    for i in [0, 7, 20]:
        for d in range(3):
            sp = s.copy(); sp[i, d] += eps
            sm = s.copy(); sm[i, d] -= eps
            g = (m.energies(sp)[0] - m.energies(sm)[0]) / (2 * eps)
            assert g == pytest.approx(h[i, d], abs=1e-7)


def test_lieb_asymmetry_and_disorder_fields():
    nb, rv, gm, zn = cpu_ref.lattice("lieb", 4)
    N = nb.shape[0]
    rng = np.random.default_rng(5)
    imp = (rng.random(N) < 0.3).astype(np.int32)
    kw = dict(J1=1.0, J2=0.2, J2d=0.05, imp=0.3, J1ic=0.5, J1ii=0.25, J2ic=0.1, J2ii=0.05)
    m = cpu_ref.Model(nb, zn, params(**kw), imp=imp, lieb=True)
    s = _random_spins(rng, 1, N)[0]
    h = m.local_fields(s)[0]
    # explicit restatement for one site
    i = 5
    J1m = [[1.0, 0.5], [0.5, 0.25]]
    J2m = [[0.2, 0.1], [0.1, 0.05]]
    exp = np.zeros(3)
    for k in range(4):
        j = nb[i, k]
        exp += J1m[imp[i]][imp[j]] * s[j]
    for k in range(4):
        j = nb[i, 4 + k]
        J = J2m[imp[i]][imp[j]] + (-1 if i & 1 else 1) * (-1 if k & 1 else 1) * 0.05
        exp += J * s[j]
    np.testing.assert_allclose(h[i], exp, rtol=1e-14)


# ------------------------------------------------------------ updates

def test_reflection_invariants():
    rng = np.random.default_rng(0)
    for _ in range(100):
        s = rng.normal(size=3); s /= np.linalg.norm(s)
        h = rng.normal(size=3) * 3
        o = cpu_ref.reflect(s, h)
        assert np.linalg.norm(o) == pytest.approx(1.0, abs=1e-14)
        assert o @ h == pytest.approx(s @ h, abs=1e-13)
        np.testing.assert_allclose(cpu_ref.reflect(o, h), s, atol=1e-14)
    s = np.array([0.0, 0.6, 0.8])
    np.testing.assert_array_equal(cpu_ref.reflect(s, np.array([1e-7, 0, 0])), s)


def test_heat_bath_distribution_is_langevin():
    rng = np.random.default_rng(7)
    beta, h = 2.0, np.array([0.3, -1.1, 0.7])
    H = np.linalg.norm(h)
    n = 40000
    chi = np.empty(n)
    for k in range(n):
        o = cpu_ref.heat_bath(beta, h, rng.random(), rng.random())
        assert abs(np.linalg.norm(o) - 1) < 1e-12
        chi[k] = o @ h / H
    bh = beta * H
    # density of chi = cos(angle to h) is proportional to exp(-bh chi)
    mean = 1 / bh - 1 / np.tanh(bh)
    assert abs(chi.mean() - mean) < 4 * chi.std() / math.sqrt(n)
    # KS test against the exact CDF
    xs = np.sort(chi)
    cdf = (1 - np.exp(-bh * (1 + xs))) / (1 - np.exp(-2 * bh))
    D = np.max(np.abs(cdf - np.arange(1, n + 1) / n))
    assert D < 1.63 / math.sqrt(n)   # 1 % level


# -------------------------------------------------------------- Ising

def _ising_4x4_exact(beta, J1=1.0):
    nb, rv, gm, zn = cpu_ref.lattice("square", 4)
    N = 16
    Es = np.zeros(1 << N)
    bonds = [(i, j) for i in range(N) for j in nb[i, :4] if i < j]
    idx = np.arange(1 << N)
    spins = 1 - 2 * ((idx[:, None] >> np.arange(N)) & 1)
    for i, j in bonds:
        Es += J1 * spins[:, i] * spins[:, j]
    w = np.exp(-beta * (Es - Es.min()))
    return (w * Es).sum() / w.sum()


def test_ising_reference_order_matches_enumeration():
    nb, rv, gm, zn = cpu_ref.lattice("square", 4)
    m = cpu_ref.Model(nb, zn, params(J1=1.0))
    temps = np.array([1.5, 2.5, 4.0])
    R = len(temps)
    rng = np.random.default_rng(11)
    s = np.zeros((R, 16, 3)); s[:, :, 0] = rng.choice([-1.0, 1.0], size=(R, 16))
    n = 40000
    cos, acc = m.mc_reference(1, temps, s, n, pt_start=n + 1, seed=9)
    for k, T in enumerate(temps):
        exact = _ising_4x4_exact(1 / T)
        est = acc[k, 0] / n
        # autocorrelated samples: allow a generous 6 sigma on the naive error
        var = acc[k, 1] / n - est ** 2
        assert abs(est - exact) < 6 * math.sqrt(var / n) * 5 + 0.05, (T, est, exact)


def test_open_chain_heisenberg_energy_is_exact():
    # classical chain, H = J sum S_i.S_i+1, <S.S'> = 1/(bJ) - coth(bJ)
    N = 64
    nb = -np.ones((N, 3), np.int32)
    for i in range(N):
        if i > 0: nb[i, 0] = i - 1
        if i < N - 1: nb[i, 1] = i + 1
    m = cpu_ref.Model(nb, (2, 0, 1), params(J1=1.0))
    temps = np.array([0.5, 1.0, 2.0])
    rng = np.random.default_rng(4)
    s = _random_spins(rng, 3, N)
    n = 3000
    cos, acc = m.mc_reference(0, temps, s, n, n_or=1, n_hb=1, pt_start=n + 1, seed=3)
    for k, T in enumerate(temps):
        b = 1 / T
        exact = (N - 1) * (1 / b - 1 / np.tanh(b))
        assert acc[k, 0] / n == pytest.approx(exact, abs=0.6), T


def test_open_chain_energy_is_exact_with_replica_exchange():
    """The same exact open-chain energies per temperature slot with the
    replica exchange on (every sweep, alternating pair parity, the 50 % cap
    and the (2, 8) pass boost): an exchange that broke detailed balance --
    a wrong sign in the exponent, the pair's temperatures or energies swapped
    -- moves configurations to the wrong temperatures and biases the slot
    energies by several units."""
    N = 64
    nb = -np.ones((N, 3), np.int32)
    for i in range(N):
        if i > 0: nb[i, 0] = i - 1
        if i < N - 1: nb[i, 1] = i + 1
    m = cpu_ref.Model(nb, (2, 0, 1), params(J1=1.0))
    temps = np.array([0.5, 0.6, 0.72, 0.86, 1.03, 1.24])
    R = len(temps)
    rng = np.random.default_rng(14)
    s = _random_spins(rng, R, N)
    st = (np.arange(R, dtype=np.int32), np.zeros(R, np.int64), np.zeros(R, np.int64),
          np.zeros(R, np.int32))
    m.mc_reference(0, temps, s, 500, n_or=1, n_hb=1, pt_start=0, seed=21, state=st)
    n = 6000
    cos, acc = m.mc_reference(0, temps, s, n, n_or=1, n_hb=1, pt_start=0, seed=22, state=st,
                              sweep0=500)
    assert st[2].sum() > 0.1 * st[1].sum()   # exchanges happen
    for k, T in enumerate(temps):
        b = 1 / T
        exact = (N - 1) * (1 / b - 1 / np.tanh(b))
        est = acc[k, 0] / n
        err = math.sqrt((acc[k, 1] / n - est ** 2) / n)
        # allowance x4 for the autocorrelation of successive sweeps
        assert abs(est - exact) < 4 * 4 * err, (T, est, exact, err)


def test_ising_enumeration_with_replica_exchange():
    nb, rv, gm, zn = cpu_ref.lattice("square", 4)
    m = cpu_ref.Model(nb, zn, params(J1=1.0))
    temps = np.array([1.5, 1.9, 2.5, 3.2])
    R = len(temps)
    rng = np.random.default_rng(12)
    s = np.zeros((R, 16, 3)); s[:, :, 0] = rng.choice([-1.0, 1.0], size=(R, 16))
    st = (np.arange(R, dtype=np.int32), np.zeros(R, np.int64), np.zeros(R, np.int64),
          np.zeros(R, np.int32))
    n = 40000
    cos, acc = m.mc_reference(1, temps, s, n, pt_start=0, seed=19, state=st)
    assert st[2].sum() > 0.1 * st[1].sum()
    for k, T in enumerate(temps):
        exact = _ising_4x4_exact(1 / T)
        est = acc[k, 0] / n
        var = acc[k, 1] / n - est ** 2
        assert abs(est - exact) < 6 * math.sqrt(var / n) * 5 + 0.05, (T, est, exact)


# ----------------------------------------------------- replica exchange

@pytest.mark.parametrize("fc", [-3.0, -1.0, -0.2, 0.0, 0.7])
def test_exchange_accepts_with_the_metropolis_probability(fc):
    """Two replicas, fixed configuration energies: over fresh attempts the
    swap is accepted with probability min(1, exp((b2 - b1)(E2 - E1)))
    (CF/mc_hm_code.cpp:3080-3095), within binomial error."""
    beta = np.array([1.0, 1.5])
    E = np.array([0.0, fc / 0.5])
    n = 40000
    acc = 0
    for t in range(n):
        st = cpu_ref.exchange(E, beta, 0, cap50=False, seed=7, sweep=t)
        acc += int(st[2][1])
        assert st[1].tolist() == [1, 1]
        assert st[2][0] == st[2][1] and st[3][0] == st[3][1] == st[2][1]
        assert st[0].tolist() == ([1, 0] if st[2][1] else [0, 1])
    p = min(1.0, math.exp(fc))
    sig = math.sqrt(max(p * (1 - p), 1e-12) / n)
    assert abs(acc / n - p) < 4 * sig + 1e-12, (fc, acc / n, p)


def test_exchange_pairs_by_parity():
    E = np.array([0.0, 0.0, 0.0, 0.0, 0.0])
    beta = np.array([1.0, 1.1, 1.2, 1.3, 1.4])
    st = cpu_ref.exchange(E, beta, 1, cap50=False)
    assert st[1].tolist() == [0, 1, 1, 1, 1]      # pairs (1, 2), (3, 4)
    st = cpu_ref.exchange(E, beta, 0, cap50=False)
    assert st[1].tolist() == [1, 1, 1, 1, 0]      # pairs (0, 1), (2, 3); slot 4 idle


def test_exchange_cap_uses_the_receivers_ratio():
    """The 50 % cap: the receiver k2 = k1 + 1 accepts only while
    (ptExc + 1) 100 / ptCnt <= 50 after counting the attempt
    (CF/mc_hm_code.cpp:3075-3099).  Three slots with equal energies (every
    draw accepted without the cap) and alternating parity: worked out by hand
    from that rule, slot 1's counters decide the pair (0, 1) and slot 2's the
    pair (1, 2), accepting at rounds 2, 3, 6, 7 of the first eight; with the
    sender's counters the pattern would start F, T."""
    E = np.zeros(3)
    beta = np.array([1.0, 1.2, 1.4])
    st = None
    pattern = []
    for r in range(8):
        before = 0 if st is None else int(st[2].sum())
        st = cpu_ref.exchange(E, beta, r % 2, cap50=True, seed=3, sweep=r, state=st)
        pattern.append(int(st[2].sum()) > before)
    assert pattern == [False, False, True, True, False, False, True, True]
    assert st[1].tolist() == [4, 8, 4] and st[2].tolist() == [2, 4, 2]


@pytest.mark.parametrize("p", [1.0, 0.8, 0.3])
def test_exchange_cap_limits_the_long_run_rate(p):
    """With the cap a pair whose Metropolis probability exceeds 1/2 is
    accepted at a long-run rate of 1/2 (exactly every second attempt at
    p = 1); below 1/2 the cap rarely binds and the rate is p."""
    beta = np.array([1.0, 2.0])
    E = np.array([0.0, math.log(p)])
    st = None
    n = 20000
    for t in range(n):
        cos = np.arange(2, dtype=np.int32)   # fresh configurations: p stays fixed
        st = cpu_ref.exchange(E, beta, 0, cap50=True, seed=11, sweep=t,
                              state=None if st is None else (cos, st[1], st[2], st[3]))
    rate = st[2][1] / n
    if p == 1.0:
        assert st[2][1] == n // 2
    else:
        want = min(p, 0.5)
        assert abs(rate - want) < 4 * math.sqrt(want * (1 - want) / n) + 2.0 / n, (p, rate)


# ---------------------------------------------------------------- SMD

def test_larmor_precession_rk4():
    # one spin, no neighbours, field h along z: dS/dt = (-h z) x S
    nb = -np.ones((1, 3), np.int32)
    h = 0.7
    m = cpu_ref.Model(nb, (1, 1, 1), params(J1=1.0, h=h))
    s0 = np.array([[0.6, 0.0, 0.8]])
    dt, T = 0.01, 501
    s, tr, q, e = m.smd_run(s0, dt, T)
    t = np.arange(T) * dt
    exact = np.stack([0.6 * np.cos(h * t), -0.6 * np.sin(h * t), 0.8 * np.ones(T)], axis=1)
    np.testing.assert_allclose(tr[:, 0, :], exact, atol=1e-10)
    # RK4 global error is O(dt^4): halving dt shrinks it ~16x
    s2, tr2, _, _ = m.smd_run(s0, 2 * dt, 251)
    err1 = np.abs(tr[::2, 0] - exact[::2]).max()
    err2 = np.abs(tr2[:, 0] - exact[::2]).max()
    assert err2 / max(err1, 1e-300) > 8


def test_smd_conserves_norm_energy_and_total_sz():
    nb, rv, gm, zn = cpu_ref.lattice("square", 6)
    m = cpu_ref.Model(nb, zn, params(J1=1.0, J2=0.6))
    rng = np.random.default_rng(8)
    s0 = _random_spins(rng, 1, 36)[0]
    s, tr, q, e = m.smd_run(s0, 0.01, 201)
    np.testing.assert_allclose(np.linalg.norm(tr, axis=-1), 1.0, atol=1e-13)
    np.testing.assert_allclose(e, e[0], rtol=1e-10)
    sz = tr[:, :, 2].sum(axis=1)
    np.testing.assert_allclose(sz, sz[0], atol=1e-6)
    assert q[0] == 1.0


def test_two_spin_dimer_relative_precession():
    # two spins coupled by J: total S is conserved and each spin precesses
    # about it with angular frequency J |S_tot|
    nb = np.array([[1, -1, -1], [0, -1, -1]], np.int32)
    J = 1.3
    m = cpu_ref.Model(nb, (1, 1, 1), params(J1=J))
    a = np.array([1.0, 0, 0]); b = np.array([0.0, 0.6, 0.8])
    s, tr, q, e = m.smd_run(np.stack([a, b]), 0.005, 401)
    tot = tr.sum(axis=1)
    np.testing.assert_allclose(tot, np.broadcast_to(tot[0], tot.shape), atol=1e-12)
    St = tot[0]
    w = J * np.linalg.norm(St)
    u = St / np.linalg.norm(St)
    t = 400 * 0.005
    # dS1/dt = J S2 x S1 = J (S_tot x S1): rotation about u by angle +w t
    par = (a @ u) * u
    perp = a - par
    exact = par + perp * np.cos(w * t) + np.cross(u, perp) * np.sin(w * t)
    np.testing.assert_allclose(tr[-1, 0], exact, atol=1e-9)


def test_gauss_seidel_and_jacobi_corrections_agree():
    nb, rv, gm, zn = cpu_ref.lattice("square", 6)
    m = cpu_ref.Model(nb, zn, params(J1=1.0, J2=0.6))
    rng = np.random.default_rng(9)
    s0 = _random_spins(rng, 1, 36)[0]
    a = m.smd_run(s0, 0.01, 101, gauss_seidel=True)[1]
    b = m.smd_run(s0, 0.01, 101, gauss_seidel=False)[1]
    np.testing.assert_allclose(a, b, atol=1e-10)


def test_dssf_matches_numpy_fft_and_parseval():
    L, T = 4, 32
    nb, rv, gm, zn = cpu_ref.lattice("square", L)
    rng = np.random.default_rng(12)
    traj = _random_spins(rng, T, L * L)
    w = cpu_ref.window("hann1", T)
    out = cpu_ref.dssf(traj, gm, L, w, T)
    # textbook: grid scatter, forward 2D FFT / sqrt(Nq), window, FFT over t / sqrt(T)
    grid = np.zeros((T, 3, L, L))
    for q, s in enumerate(gm):
        grid[:, :, q // L, q % L] = traj[:, s, :]
    Sq = np.fft.fft2(grid, axes=(2, 3)) / L
    X = np.fft.fft(w[:, None, None, None] * Sq, axis=0) / np.sqrt(T)
    ref = np.abs(X) ** 2                     # (T, 3, L, L)
    ref = ref.transpose(1, 2, 3, 0).reshape(3, L * L, T)
    np.testing.assert_allclose(out, ref, rtol=1e-10, atol=1e-12)
    # Parseval: sum_w |X|^2 = sum_t |w S(q,t)|^2
    lhs = out.sum(axis=2)
    rhs = (np.abs(w[:, None, None, None] * Sq) ** 2).sum(axis=0).reshape(3, L * L)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-10)


@pytest.mark.parametrize("geom,L", [("square", 4), ("kagome", 3), ("hexagn", 3)])
def test_dssf_fft_form_equals_direct_form(geom, L):
    """cpu_ref.dssf_fft (numpy FFTs, for long trajectories) equals orc_dssf
    (direct time sums) on grids with empty points, a window and n_omega < T."""
    nb, rv, gm, zn = cpu_ref.lattice(geom, L)
    G = int(round(math.sqrt(gm.size)))
    rng = np.random.default_rng(22)
    T = 48
    traj = _random_spins(rng, T, nb.shape[0])
    w = cpu_ref.window("black", T)
    a = cpu_ref.dssf(traj, gm, G, w, 20)
    b = cpu_ref.dssf_fft(traj, gm, G, w, 20)
    np.testing.assert_allclose(b, a, rtol=1e-10, atol=1e-12 * a.max())


def test_windows_match_textbook_definitions():
    n = 257
    np.testing.assert_allclose(cpu_ref.window("nowin", n), np.ones(n))
    np.testing.assert_allclose(cpu_ref.window("hann1", n), sw.hann(n, sym=True), atol=1e-15)
    np.testing.assert_allclose(cpu_ref.window("hann2", n), sw.hann(n, sym=True) ** 2, atol=1e-15)
    np.testing.assert_allclose(cpu_ref.window("black", n), sw.blackmanharris(n, sym=True), atol=1e-15)
    np.testing.assert_allclose(cpu_ref.window("ntall", n), sw.nuttall(n, sym=True), atol=1e-15)
    np.testing.assert_allclose(cpu_ref.window("flatw", n), sw.flattop(n, sym=True), atol=1e-9)
    x = np.linspace(-1, 1, n)
    np.testing.assert_allclose(cpu_ref.window("gauss", n), np.exp(-80 * x ** 2), atol=1e-15)
    np.testing.assert_allclose(cpu_ref.window("sincw", n), np.sinc(2 * x), atol=1e-14)
    lz = np.where(np.abs(0.5 * x) > 0.5, 0.0, np.sinc(x))
    np.testing.assert_allclose(cpu_ref.window("lancz", n), lz, atol=1e-14)


def test_order_parameters_on_ordered_states():
    # perfect states give the textbook values of the reference's comments
    L = 6
    nb, rv, gm, zn = cpu_ref.lattice("square", L)
    ii, jj = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    neel = np.zeros((L * L, 3)); neel[:, 2] = np.cos(np.pi * (ii + jj)).ravel()
    o = cpu_ref.observables("square", L, nb, zn, 0, neel)
    assert o[0] == pytest.approx(0, abs=1e-14)
    assert o[2] == pytest.approx(1.0, abs=1e-14)            # staggered magnetisation
    assert o[5] == pytest.approx(1.0, abs=1e-14)            # diagonal stripe = Neel on square
    hst = np.zeros((L * L, 3)); hst[:, 2] = np.cos(np.pi * jj).ravel()
    o = cpu_ref.observables("square", L, nb, zn, 0, hst)
    assert o[3] == pytest.approx(1.0, abs=1e-14) and o[4] == pytest.approx(0.0, abs=1e-7)
    assert abs(o[7]) == pytest.approx(1.0, abs=1e-14)       # Z2 parameter +-1 on stripes
    ferro = np.zeros((L * L, 3)); ferro[:, 0] = 1.0
    o = cpu_ref.observables("square", L, nb, zn, 0, ferro)
    assert o[0] == pytest.approx(1.0) and o[2] == pytest.approx(0.0, abs=1e-7)
    # triangular: 120 degree state -> Psi = sqrt(6) with spins in the (z, y) plane
    nb, rv, gm, zn = cpu_ref.lattice("triang", L)
    th = (4 * np.pi / 3) * (ii + 0.5 * jj)
    s = np.zeros((L * L, 3)); s[:, 2] = np.cos(th).ravel(); s[:, 1] = np.sin(th).ravel()
    o = cpu_ref.observables("triang", L, nb, zn, 1, s)
    assert o[7] == pytest.approx(np.sqrt(6.0), rel=1e-12)
    assert o[2] == pytest.approx(1.0, rel=1e-12)
    # 2-up 1-down collinear state (make_2U1DPhaseList): Psi = 8 / sqrt(6), the
    # value quoted in the reference's cplxC3Parameter comment
    n = (2 * ii + jj + 2) % 3
    s = np.zeros((L * L, 3)); s[:, 2] = np.cos(n * np.pi).ravel()
    o = cpu_ref.observables("triang", L, nb, zn, 1, s)
    assert o[7] == pytest.approx(8 / np.sqrt(6.0), rel=1e-12)
