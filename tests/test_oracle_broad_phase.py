"""Pins the oracle's linked-cell broad phase (LinkedCell.cpp:958-1103 with the
periodic clones of :1765-2398) against O(N^2) all-pairs with the minimum-image
convention, and its binning (Cell::GetCell, Cell.cpp:172-184) against the cell
bounds."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from tests.helpers import min_image_pairs, pair_set


def jammed(n_side, a, rng, rmin=0.97e-3, rmax=1.03e-3, jit=4e-5):
    idx = np.stack(np.meshgrid(range(n_side), range(n_side), range(n_side), indexing="ij"), -1).reshape(-1, 3)
    x = (idx + 0.5) * a + rng.uniform(-jit, jit, (len(idx), 3))
    r = rng.uniform(rmin, rmax, len(idx))
    return x, r


@pytest.mark.parametrize("periodic", [False, True])
def test_pairs_match_all_pairs(periodic):
    rng = np.random.default_rng(5)
    m, a = 8, 2.0e-3
    x, r = jammed(m, a, rng)
    L = m * a
    o = Oracle((0, 0, 0), (L, L, L), periodic=(int(periodic),) * 3, gravity=(0, 0, 0), dt=1e-5)
    o.contact_model(0, 0, 1e4, 0.8, 0.3)
    o.add_spheres(r, 2500.0, 5e-5, 0, x)
    o.init()
    got = pair_set(o.pairs())
    ref = min_image_pairs(x, r, L=(L, L, L) if periodic else None)
    assert len(ref) > 500
    assert got == ref
    # every contact is seen from both sides
    assert all((j, i) in got for (i, j) in got)


def test_periodic_pairs_follow_wrapping_motion():
    # particles cross the periodic faces for many steps; the contact set must
    # keep matching all-pairs minimum image on the masters' positions
    rng = np.random.default_rng(9)
    m, a = 6, 2.05e-3
    x, r = jammed(m, a, rng, jit=2e-5)
    L = m * a
    v = rng.normal(0, 0.3, x.shape)
    o = Oracle((0, 0, 0), (L, L, L), periodic=(1, 1, 1), gravity=(0, 0, 0), dt=1e-5)
    o.contact_model(0, 0, 1e4, 0.8, 0.3)
    o.add_spheres(r, 2500.0, 5e-5, 0, x, v)
    o.init()
    for k in range(20):
        o.step(10)
        s = o.state()
        assert np.all(s["x"] >= 0) and np.all(s["x"] < L)
        ref = min_image_pairs(s["x"], r, L=(L, L, L))
        assert pair_set(o.pairs()) == ref
    # total momentum is conserved by the contact forces
    s = o.state()
    mass, _, _ = o.masses()
    p0, p1 = (mass[:, None] * v).sum(axis=0), (mass[:, None] * s["v"]).sum(axis=0)
    assert np.allclose(p1, p0, rtol=0, atol=1e-12 * np.abs(mass[:, None] * v).sum())


def test_binning_matches_cell_bounds():
    rng = np.random.default_rng(2)
    n = 3000
    x = rng.uniform(0, [0.02, 0.03, 0.05], (n, 3))
    o = Oracle((0, 0, 0), (0.02, 0.03, 0.05), gravity=(0, 0, 0), cellsize_factor=130.0)
    o.contact_model(0, 0, 1e4, 0.8, 0.3)
    o.add_spheres(1e-5, 2500.0, 1e-6, 0, x)
    o.init()
    dims, lo, edge = o.grid()
    # cell size 2 * Rmax * factor, n = int((L + EPS) / size), edge = L / n
    for a, L in enumerate((0.02, 0.03, 0.05)):
        assert dims[a] == int((L + 1e-10) / (2 * (130.0 * 1e-5)))
        assert edge[a] == L / dims[a]
    c = o.cells()
    assert np.all(lo + c * edge <= x + 1e-15) and np.all(x < lo + (c + 1) * edge + 1e-15)
    assert np.all(c >= 0) and np.all(c < np.asarray(dims))
