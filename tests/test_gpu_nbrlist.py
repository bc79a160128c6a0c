"""The per-step neighbour candidate list (k_detect, then the force sweep over
the listed candidates) against the force sweep that searches the 27 cells
itself: same candidates in the same order, so every state array is
bit-identical after many steps, for spheres with walls, periodic gases,
dense packings (many candidates per particle, and with a capacity so small
that most particles overflow the list and are swept in full) and polyhedra.
"""
import numpy as np
import pytest

from pacific_amd import MIXED, workloads
from tests.helpers import make_gpu, pair_set

pytestmark = pytest.mark.gpu

CASES = {
    "settling": (lambda: workloads.settling_box(1000), 300),
    "gas": (lambda: workloads.periodic_gas(27000, sigma=0.6), 300),
    "dense_periodic": (lambda: workloads.dense_pack(side=8, periodic=True, gravity=(0.0, 0.0, 0.0)), 200),
    "dense_walls": (lambda: workloads.dense_pack(side=8), 200),
    "polyhedra": (lambda: workloads.polyhedra_box((5, 5, 4), seed=9, spheres=True), 300),
}


@pytest.mark.parametrize("precision", [0, MIXED])
@pytest.mark.parametrize("cap", [32, 3])
@pytest.mark.parametrize("case", sorted(CASES))
def test_list_equals_cell_sweep(gpu, case, cap, precision):
    make, nsteps = CASES[case]
    wl = make()
    a = make_gpu(wl, precision=precision, nbr=0)
    b = make_gpu(wl, precision=precision, nbr=cap)
    a.step(nsteps)
    b.step(nsteps)
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "w", "wb", "q", "F", "T"):
        assert np.array_equal(sa[k], sb[k]), k
    pa, pb = pair_set(a.pairs()), pair_set(b.pairs())
    assert pa == pb
    if case.startswith("dense"):
        assert len(pa) > 3 * a.n  # several candidates per particle: cap 3 overflows
