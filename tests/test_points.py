"""Irregular graphs from point sets (host code, CPU only): the seeded
Poisson-disk generator and the distance-shell neighbour tables, checked
against scipy's periodic k-d tree and a brute-force numpy restatement."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import classical_mc_and_smd_for_2d_heisenberg_models_amd as H


def test_poisson_disk_is_a_periodic_poisson_disk_set():
    box, r = 30.0, 1.0
    p = H.poisson_disk(box, r, seed=5)
    assert np.all((p >= 0) & (p < box))
    tree = cKDTree(p, boxsize=box)
    # minimum distance
    d, _ = tree.query(p, k=2)
    assert d[:, 1].min() >= r * (1 - 1e-12)
    # (near-)maximal: no probe point is farther than 2 r from every sample
    rng = np.random.default_rng(0)
    probe = rng.random((20000, 2)) * box
    dp, _ = tree.query(probe, k=1)
    assert dp.max() < 2 * r
    # deterministic in the seed, different across seeds
    np.testing.assert_array_equal(p, H.poisson_disk(box, r, seed=5))
    assert len(H.poisson_disk(box, r, seed=6)) != len(p) or not np.array_equal(H.poisson_disk(box, r, seed=6), p)
    # density of a Bridson sampling: roughly 0.6-0.7 points per r^2
    assert 0.55 < len(p) / box ** 2 < 0.75


def test_poisson_disk_respects_max_points_and_rejects_bad_input():
    p = H.poisson_disk(30.0, 1.0, seed=1, max_points=50)
    assert len(p) == 50
    with pytest.raises(H.HSMDError):
        H.poisson_disk(3.0, 1.0)


def test_from_points_tables_match_brute_force_shells():
    box, r = 12.0, 1.0
    rs = (1.15, 1.45, 1.75)
    p = H.poisson_disk(box, r, seed=3)
    lat = H.Lattice.from_points(p, box, rs, H.Couplings(J1=1.0, J2=0.4, JX=0.1))
    n = len(p)
    d = p[:, None, :] - p[None, :, :]
    d -= box * np.round(d / box)
    dist = np.sqrt((d ** 2).sum(-1))
    shells = []
    for s in range(3):
        lo = 0.0 if s == 0 else rs[s - 1]
        rows = []
        for i in range(n):
            js = [j for j in range(n) if j != i and lo <= dist[i, j] < rs[s]]
            js.sort(key=lambda j: (dist[i, j], j))
            rows.append(js)
        shells.append(rows)
    zn = [max(len(r_) for r_ in rows) for rows in shells]
    assert (lat.zn1, lat.zn2, lat.znx) == tuple(zn)
    ref = -np.ones((n, sum(zn)), np.int32)
    for i in range(n):
        off = 0
        for s in range(3):
            ref[i, off:off + len(shells[s][i])] = shells[s][i]
            off += zn[s]
    np.testing.assert_array_equal(lat.tables(), ref)
    np.testing.assert_allclose(lat.positions(), p)
    # a tabulated graph: the even integer part of sqrt(N) sets the q grid
    assert lat.L == (int(np.sqrt(n)) // 2) * 2
