"""Contact statistics of the time loop (grains_gpu_contact_stats, the
counterpart of AppCollision::addToContactsFeatures, AppCollision.cpp:159-189)
against the oracle's, which follows the reference's evaluation of each pair
once per pair of existing images (masters and their periodic clones):

* dense packings with walls and triply periodic, an inclined plane, with the
  Verlet list and with the per-step list: the same number of overlaps and
  closed-form evaluations, maximum and mean overlap within 1e-9 (the
  trajectories agree to rounding; the mean is a sum over a count here and a
  running mean in the reference), the time of the maximum;
* polyhedra: GJK evaluations counted alike, mean iterations within 1e-3;
* statistics are off by default and cleared when turned on.
"""
import numpy as np
import pytest

from pacific_amd import workloads
from tests.helpers import make_gpu, make_oracle

pytestmark = pytest.mark.gpu

CASES = {
    "dense_walls": (lambda: workloads.dense_pack(side=6), 200),
    "dense_periodic": (lambda: workloads.dense_pack(side=6, periodic=True, gravity=(0.0, 0.0, 0.0)), 200),
    "inclined": (lambda: workloads.inclined_plane(3000), 400),
}


@pytest.mark.parametrize("skin", [None, 0.0])
@pytest.mark.parametrize("case", sorted(CASES))
def test_contact_statistics_match_oracle(gpu, case, skin):
    make, n = CASES[case]
    wl = make()
    o = make_oracle(wl)
    g = make_gpu(wl, skin=skin, stats=True)
    for _ in range(4):
        o.step(n // 4)
        g.step(n // 4)
    so, sg = o.stats(), g.contact_stats()
    assert so["n_overlaps"] > 0
    assert sg["n_overlaps"] == so["n_overlaps"]
    assert sg["n_evaluations"] == so["n_evaluations"]
    assert sg["overlap_max"] == pytest.approx(so["overlap_max"], rel=1e-9)
    assert sg["overlap_mean"] == pytest.approx(so["overlap_mean"], rel=1e-9)
    assert abs(sg["time_overlap_max"] - so["time_overlap_max"]) <= 1e-12 + 0.5 * wl["dt"]
    assert sg["gjk_iter_mean"] == 1.0


def test_gjk_iteration_statistics(gpu):
    wl = workloads.polyhedra_box((4, 4, 3), seed=9, spheres=True, sigma=0.3)
    o = make_oracle(wl)
    g = make_gpu(wl, stats=True)
    o.step(300)
    g.step(300)
    so, sg = o.stats(), g.contact_stats()
    assert so["n_evaluations"] > 100
    assert sg["n_evaluations"] == so["n_evaluations"]
    assert sg["n_overlaps"] == so["n_overlaps"]
    assert sg["gjk_iter_mean"] == pytest.approx(so["gjk_iter_mean"], rel=1e-3)
    assert sg["gjk_iter_mean"] > 1.0
    assert sg["overlap_max"] == pytest.approx(so["overlap_max"], rel=1e-6)


def test_statistics_off_by_default_and_cleared(gpu):
    wl = workloads.dense_pack(side=4)
    g = make_gpu(wl)
    g.step(20)
    assert g.contact_stats()["n_overlaps"] == 0
    g.contact_stats(True)
    g.step(10)
    a = g.contact_stats()
    assert a["n_overlaps"] > 0
    g.contact_stats(True)
    assert g.contact_stats()["n_overlaps"] == 0
    g.compute_forces()  # outside the time loop: not counted
    assert g.contact_stats()["n_overlaps"] == 0
