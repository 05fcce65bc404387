"""Pins the oracle's MC measurement restatements (spin stiffness, static S(q)
on the DFT grid and by direct sums, the key points and correlation lengths)
against derivatives of the oracle's energy, numpy's FFT and ordered states
whose spectra are known.  CPU only."""
import math

import numpy as np
import pytest

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
