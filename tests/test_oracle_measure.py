"""Pins the oracle's MC measurement restatements (spin stiffness, static S(q)
on the DFT grid and by direct sums, the key points and correlation lengths)
against derivatives of the oracle's energy, numpy's FFT and ordered states
whose spectra are known.  CPU only."""
import math
import os

import numpy as np
import pytest

import classical_mc_and_smd_for_2d_heisenberg_models_amd as H
import cpu_ref
from test_oracle import _random_spins, params

C60, S60 = math.cos(math.pi / 3), math.sin(math.pi / 3)
DIRS = [(1.0, 0.0), (C60, S60), (-C60, S60)]


def _twist(spins, rvec, a, phi):
    """Rotate every spin about z by phi * (a . r_i)."""
    th = phi * (a[0] * rvec[:, 0] + a[1] * rvec[:, 1])
    c, s = np.cos(th), np.sin(th)
    out = spins.copy()
    out[:, 0] = c * spins[:, 0] - s * spins[:, 1]
    out[:, 1] = s * spins[:, 0] + c * spins[:, 1]
    return out


@pytest.mark.parametrize("geom,L,kw,dis", [
    ("square", 6, dict(J1=1.0, J2=0.6, l1=0.8), False),
    ("triang", 6, dict(J1=1.0, J2=0.12, JX=0.05, l1=0.9, l2=0.9), False),
    ("square", 6, dict(J1=1.0, J2=0.3, imp=0.3, J1ic=0.5, J1ii=0.25, J2ic=0.1, J2ii=0.05), True),
])
def test_spin_stiffness_is_the_twist_derivative_of_the_energy(geom, L, kw, dis):
    # Twisting S_i about z by phi (a . r_i) changes each bond's xy product to
    # d cos(phi C) + v sin(phi C), C = a . (r_i - r_j): E'(0) = sum J C v and
    # E''(0) = -sum J C^2 d, independent of the stiffness code.
    nb, rv, gm, zn = cpu_ref.lattice(geom, L)
    N = nb.shape[0]
    rng = np.random.default_rng(11)
    imp = (rng.random(N) < 0.3).astype(np.int32) if dis else None
    m = cpu_ref.Model(nb, zn, params(**kw), imp=imp)
    s = _random_spins(rng, 1, N)[0]
    st = m.spin_stiffness(rv, s)[0]
    phi = 1e-3
    for k, a in enumerate(DIRS):
        e = [m.energies(_twist(s, rv, a, f))[0] for f in (-phi, 0.0, phi)]
        d1 = (e[2] - e[0]) / (2 * phi)
        d2 = (e[2] - 2 * e[1] + e[0]) / phi ** 2
        scale = abs(st[k]) + 1.0
        assert d1 == pytest.approx(st[3 + k], abs=1e-5 * (abs(st[3 + k]) + scale))
        assert -d2 == pytest.approx(st[k], abs=2e-4 * scale)


@pytest.mark.parametrize("geom,L", [("square", 6), ("kagome", 4), ("hexagn", 4)])
def test_sq_grid_matches_numpy_fft2(geom, L):
    nb, rv, gm, zn = cpu_ref.lattice(geom, L)
    N = nb.shape[0]
    G = int(round(math.sqrt(gm.size)))
    rng = np.random.default_rng(3)
    s = _random_spins(rng, 2, N)
    out = cpu_ref.sq_grid(gm, G, s)
    for r in range(2):
        for d in range(3):
            grid = np.zeros(G * G)
            grid[gm >= 0] = s[r, gm[gm >= 0], d]
            ref = np.fft.fft2(grid.reshape(G, G)).reshape(-1) / math.sqrt(G * G)
            np.testing.assert_allclose(out[r, d], ref, atol=1e-12)


def test_sq_direct_equals_fft_on_integer_square_positions():
    # square lattice sites sit at integer (x, y) = (i, j): with dq = 2 pi / L
    # the direct sum at (nx, ny) is the FFT at (nx mod L, ny mod L)
    L = 6
    nb, rv, gm, zn = cpu_ref.lattice("square", L, pbc=False)
    N = L * L
    rng = np.random.default_rng(4)
    s = _random_spins(rng, 1, N)
    out = cpu_ref.sq_direct(rv, L, N, 2 * math.pi / L, s)
    for d in range(3):
        grid = np.zeros((L, L))
        ix = np.rint(rv[:, 0]).astype(int)
        iy = np.rint(rv[:, 1]).astype(int)
        grid[ix, iy] = s[0, :, d]
        F = np.fft.fft2(grid) / math.sqrt(N)
        np.testing.assert_allclose(out[0, d].reshape(L, L).T, F, atol=1e-12)
    # an Nq longer than L^2 keeps zeros beyond L^2
    out2 = cpu_ref.sq_direct(rv, L, N + 10, 2 * math.pi / L, s)
    assert np.all(out2[0, :, N:] == 0)


def _square_state(L, kind):
    ix, iy = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    if kind == "ferro":
        sgn = np.ones((L, L))
    elif kind == "neel":
        sgn = (-1.0) ** (ix + iy)
    elif kind == "stripe_i":
        sgn = (-1.0) ** ix
    else:
        sgn = (-1.0) ** iy
    s = np.zeros((L * L, 3))
    s[:, 0] = sgn.reshape(-1)
    return s


@pytest.mark.parametrize("kind,point", [("ferro", 3), ("neel", 0), ("stripe_i", 1), ("stripe_j", 2)])
def test_corr_len_key_points_on_ordered_square_states(kind, point):
    # ferromagnet -> Gamma, Neel -> M (zone corner), stripes -> X / Y
    L = 8
    nb, rv, gm, zn = cpu_ref.lattice("square", L)
    N = L * L
    s = _square_state(L, kind)
    sq = cpu_ref.sq_grid(gm, L, s)[0]
    dq = 2 * math.pi / L
    pk, cl = cpu_ref.corr_len(sq, L, False, dq)
    assert pk[0, point] == pytest.approx(N / 2, rel=1e-12)          # xy
    assert pk[3, point] == pytest.approx(N / 3, rel=1e-12)          # tt
    assert pk[2, point] == pytest.approx(1e-16, rel=1e-6)           # zz empty
    for other in range(5):
        if other != point:
            assert pk[0, other] < 1e-10
    assert cl[0, point] == pytest.approx(math.sqrt(pk[0, point] / 1e-16 - 1) / dq, rel=1e-6)
    assert cl[2, point] == 0.0


def test_corr_len_k_point_on_triangular_120_state():
    # 120-degree state, angle 2 pi (i - j) / 3 on the grid: peaks at the K
    # points with |S_x|^2 = |S_y|^2 = N / 4
    L = 6
    nb, rv, gm, zn = cpu_ref.lattice("triang", L)
    N = L * L
    s = np.zeros((N, 3))
    for n in range(L * L):
        i, j = divmod(n, L)
        k = gm[n]
        th = 2 * math.pi * (i - j) / 3
        s[k] = (math.cos(th), math.sin(th), 0.0)
    sq = cpu_ref.sq_grid(gm, L, s)[0]
    pk, cl = cpu_ref.corr_len(sq, L, True, 2 * math.pi / L)
    assert pk[0, 0] == pytest.approx(N / 4, rel=1e-12)
    assert pk[3, 0] == pytest.approx(N / 6, rel=1e-12)
    assert pk[0, 3] < 1e-10            # Gamma empty


def test_dssf_direct_equals_grid_dssf_on_square_lattice():
    # square sites at integer positions: direct sums at q = 2 pi (a, b) / L
    # reproduce the grid DSSF at FFT index (a, b) (N = Nq), with a window
    L, T, no = 4, 16, 16
    nb, rv, gm, zn = cpu_ref.lattice("square", L)
    N = L * L
    rng = np.random.default_rng(8)
    traj = _random_spins(rng, T, N)
    w = cpu_ref.window("hann1", T)
    ref = cpu_ref.dssf(traj, gm, L, w, no)
    # grid point k = a * L + b of the oracle grid DSSF: the site at grid (a', b')
    # has position (a', b'), so q = 2 pi / L (a, b)
    q = np.array([[2 * math.pi * (k // L) / L, 2 * math.pi * (k % L) / L] for k in range(N)])
    got = cpu_ref.dssf_direct(traj, rv, q, w, no)
    np.testing.assert_allclose(got, ref, atol=1e-12)
    # a masked site is removed from the sums
    mask = np.zeros(N, np.int32)
    mask[3] = 1
    t2 = traj.copy()
    t2[:, 3] = 0.0
    np.testing.assert_allclose(cpu_ref.dssf_direct(traj, rv, q, w, no, mask=mask),
                               cpu_ref.dssf_direct(t2, rv, q, w, no), atol=1e-12)


def test_dssf_direct_fft_form_equals_direct_loops():
    """cpu_ref.dssf_direct_fft (FFT in time, one matrix product in space)
    equals orc_dssf_direct on an irregular point set with a mask and a
    window, at arbitrary wavevectors."""
    rng = np.random.default_rng(9)
    N, T, no = 37, 40, 24
    rv = rng.random((N, 2)) * 6.0
    traj = _random_spins(rng, T, N)
    q = rng.normal(size=(11, 2)) * 2.0
    w = cpu_ref.window("ntall", T)
    mask = (rng.random(N) < 0.2).astype(np.int32)
    a = cpu_ref.dssf_direct(traj, rv, q, w, no, mask=mask)
    b = cpu_ref.dssf_direct_fft(traj, rv, q, w, no, mask=mask)
    np.testing.assert_allclose(b, a, rtol=1e-10, atol=1e-12 * a.max())


# ------------------------------------------------ quasicrystal observables

def _qc_fixture(pbc, n):
    d = os.path.join(os.path.dirname(__file__), "golden", "qc_%s_%d" % (pbc, n))
    rd = lambda k, dt: np.fromfile(os.path.join(d, "%s_AppxtSz(%d).bin" % (k, n)), dtype=dt)
    nb = rd("nborsList", "<i4").reshape(n, 19)
    rv = rd("sitesList", "<f8").reshape(n, 2)
    st = np.stack([rd(k, "<f8").reshape(n, 3) for k in
                   ["stateNeel0", "stateStrp1", "stateStrp2", "stateStrp3", "stateStrp4"]])
    return nb, rv, st


@pytest.mark.parametrize("pbc,n", [("with_PBC", 41), ("no_PBC", 239)])
def test_qc_projections_of_the_reference_states(pbc, n):
    # each +-1 reference state projects on itself with 1; a saturated stripe
    # projection zeroes the other three (the cut of get_qctStrpPrjt)
    nb, rv, st = _qc_fixture(pbc, n)
    v = cpu_ref.qc_values(nb, 8, st, st, rv, n, 1.0, np.zeros((8, 2), np.int32))
    assert v[0, 0] == pytest.approx(1.0, abs=1e-14)
    for i in range(4):
        assert v[1 + i, 1 + i] == pytest.approx(1.0, abs=1e-14)
        assert np.all(np.delete(v[1 + i, 1:5], i) == 0.0)
    # cross projections of unsaturated states are plain overlaps / N
    assert v[0, 1] == pytest.approx(np.sum(st[0] * st[1]) / n, abs=1e-14)


def _star():
    # site 0 with eight clock-ordered shell-1 neighbours 1..8; each of them
    # has site 0 in slot 0
    n = 9
    nb = -np.ones((n, 19), np.int32)
    nb[0, :8] = np.arange(1, 9)
    nb[1:, 0] = 0
    rv = np.zeros((n, 2))
    return nb, rv


def test_qc_stripe_vector_and_frustration_closed_forms():
    nb, rv = _star()
    s = np.zeros((9, 3))
    s[0, 0] = 1.0
    # slots 0, 1, 4, 5 up, slots 2, 3, 6, 7 down: site 0 gives (rx, ry) = (1, 0);
    # each outer site gives (+-1, -+1) clipped to length 1, cancelling in pairs
    sgn = np.array([1, 1, -1, -1, 1, 1, -1, -1], float)
    s[1:, 0] = sgn
    v = cpu_ref.qc_values(nb, 8, s, np.zeros((5, 9, 3)), rv, 9, 1.0, np.zeros((8, 2), np.int32))
    assert v[0, 5] == pytest.approx(1.0 / 9.0, abs=1e-15)
    # N1fpar: site 0 sees a zero average, outer sites (1 - s0.s_j) / 2 each
    assert v[0, 6] == pytest.approx((0.5 + np.sum(0.5 * (1 - sgn))) / 9.0, abs=1e-15)
    # staggered star: every bond antiparallel -> N1fpar = 1; ferro -> 0
    s[1:, 0] = -1.0
    v = cpu_ref.qc_values(nb, 8, s, np.zeros((5, 9, 3)), rv, 9, 1.0, np.zeros((8, 2), np.int32))
    assert v[0, 6] == pytest.approx(1.0, abs=1e-15)
    s[1:, 0] = 1.0
    v = cpu_ref.qc_values(nb, 8, s, np.zeros((5, 9, 3)), rv, 9, 1.0, np.zeros((8, 2), np.int32))
    assert v[0, 6] == pytest.approx(0.0, abs=1e-15)
    # ferro star: site 0 has vec1 = vec2 = 0; each outer site has its one
    # neighbour in slot 0, (rx, ry) = (1, -1) clipped to (1, -1)/sqrt 2:
    # |sum| / 9 = 8 / 9
    assert v[0, 5] == pytest.approx(8.0 / 9.0, abs=1e-15)


def test_qc_hotspots_and_amplitudes_of_a_plane_wave():
    # periodic 14 x 14 grid, dq = 2 pi / 14: a spiral at q0 = dq (4, 2) has
    # |S(q)|^2 only at +-q0 (discrete orthogonality), so the arc search
    # (2 < |q| < 5, 0 < arg q < pi/4) finds q0 and the eight images follow;
    # SqPeaks[0] = sqrt(0.5 / Nq (N^2/6 + N^2/6)), the other three vanish
    Lg = 14
    x, y = np.meshgrid(np.arange(Lg), np.arange(Lg), indexing="xy")
    rv = np.stack([x.ravel(), y.ravel()], 1).astype(float)
    n = Lg * Lg
    dq = 2 * math.pi / Lg
    th = dq * (4 * rv[:, 0] + 2 * rv[:, 1])
    s = np.stack([np.cos(th), np.sin(th), np.zeros(n)], 1)
    pk = cpu_ref.qc_peaks(s, rv, Lg, dq)
    want = [(4, 2), (-4, -2), (2, 4), (-2, -4), (-4, 2), (2, -4), (4, -2), (-2, 4)]
    assert [tuple(p) for p in pk.tolist()] == want
    nb = -np.ones((n, 19), np.int32)
    v = cpu_ref.qc_values(nb, 8, s, np.zeros((5, n, 3)), rv, n, dq, pk)
    assert v[0, 7] == pytest.approx(math.sqrt(n * n / (6.0 * n)), rel=1e-12)
    np.testing.assert_allclose(v[0, 8:11], 0.0, atol=1e-10)


def test_qc_hotspots_of_the_library_match_the_oracle():
    # host-side search of the library (hsmd_lattice_set_qc_states) against
    # the oracle on both fixtures and both reference choices; the 41-site
    # fixture has no grid point in the arc (hotspots at q = 0)
    for pbc, n in [("with_PBC", 41), ("no_PBC", 239)]:
        nb, rv, st = _qc_fixture(pbc, n)
        lat = H.Lattice.from_tables(nb, 8, 3, 8, rv, H.Couplings(J1=1.0, J2=0.4, JX=0.1))
        dq = 8 * math.pi / lat.Gsz
        for neel in (True, False):
            got = lat.set_qc_states(st[0], st[1:], neel_phase=neel)
            ref = cpu_ref.qc_peaks(st[0] if neel else st[1], rv, lat.L, dq)
            np.testing.assert_array_equal(got, ref)
            if n == 41:
                assert np.all(got == 0)
            else:
                assert np.any(got != 0)
