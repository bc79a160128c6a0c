"""Neighbour candidate lists against the force sweep that searches the 27
cells itself.

* per-step list (k_detect after every binning, skin 0): same candidates in
  the same order, so every state array is bit-identical after many steps,
  for spheres with walls, periodic gases, dense packings (also with a
  capacity so small that most particles overflow the list and are searched
  in their cells) and polyhedra;
* Verlet list (skin > 0, rebuilt only once a particle has moved more than
  half the skin; each step filters the listed pairs down to its contacts,
  then the force sweep visits those): with velocity-independent contact
  forces (en = 1, no friction) the torsor a step leaves is that of a force
  evaluation at the same state, bit for bit; at any step, the contacts
  found from the list -- built
  several steps earlier -- are the contacts of a fresh cell search of the
  same state, with the same per-contact arithmetic: equal contact records,
  forces and torques within 1e-12 relative (the contact sums run in the
  list's order), and the list is reused over many steps.
"""
import numpy as np
import pytest

from pacific_amd import MIXED, workloads
from tests.helpers import contact_scale, make_gpu, pair_set, rel_force_error

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
    b = make_gpu(wl, precision=precision, nbr=cap, skin=0.0)
    a.step(nsteps)
    b.step(nsteps)
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "w", "wb", "q", "F", "T"):
        assert np.array_equal(sa[k], sb[k]), k
    pa, pb = pair_set(a.pairs()), pair_set(b.pairs())
    assert pa == pb
    if case.startswith("dense"):
        assert len(pa) > 3 * a.n  # several candidates per particle: cap 3 overflows


CASES["inclined"] = (lambda: workloads.inclined_plane(5000), 300)
VERLET_CASES = ["inclined", "gas", "dense_periodic", "dense_walls"]


@pytest.mark.parametrize("precision", [0, MIXED])
@pytest.mark.parametrize("cap", [32, 3])
@pytest.mark.parametrize("case", VERLET_CASES)
def test_verlet_contacts_equal_cell_search(gpu, case, cap, precision):
    make, nsteps = CASES[case]
    wl = make()
    v = make_gpu(wl, precision=precision, nbr=cap)  # default skin: Verlet list
    c = make_gpu(wl, precision=precision, nbr=0)
    skin, _ = v.neighbor_stats()
    assert skin > 0
    chunk = max(nsteps // 4, 1)
    npairs = 0
    for _ in range(4):
        v.step(chunk)
        sv = v.state()
        # forces of the current state from the list as it stands (built at
        # some earlier step) and from a fresh cell search of the same state
        v.compute_forces()
        fv = v.state()
        c.set_state(x=sv["x"], v=sv["v"], w=sv["w"], q=sv["q"], recompute=True)
        c.compute_forces()
        fc = c.state()
        pv, pc = pair_set(v.pairs()), pair_set(c.pairs())
        assert pv == pc
        npairs += len(pv)
        if precision == 0:
            _, _, wgt = c.masses()
            scale = contact_scale(fc, wgt)
            assert rel_force_error(fv["F"], fc["F"], scale) <= 1e-12
            assert rel_force_error(fv["T"], fc["T"], scale * 1e-3) <= 1e-12
        else:
            # FP32 contact sums in another order: rounding of the contact
            # forces, which may cancel in a particle's net force -- relative
            # to the sum of the magnitudes of the particle's contact forces
            fs = c.force_sums()
            touch = fs > 0
            dF = np.linalg.norm(fv["F"] - fc["F"], axis=1)
            dT = np.linalg.norm(fv["T"] - fc["T"], axis=1)
            assert np.all(dF[~touch] == 0) and np.all(dT[~touch] == 0)
            assert np.max(dF[touch] / fs[touch]) <= 1e-6
            assert np.max(dT[touch] / (fs[touch] * 1.2e-3)) <= 1e-6
    _, builds = v.neighbor_stats()
    assert 1 <= builds < 4 * chunk  # the list outlives single steps
    assert npairs > 0


def test_verlet_rebuilds_track_motion(gpu):
    """A fast periodic gas rebuilds its list often, a packing at rest rarely;
    either way the contact records match a fresh cell search step after
    step."""
    fast = make_gpu(workloads.periodic_gas(27000, sigma=0.6))
    slow = make_gpu(workloads.dense_pack(side=8, sigma=1e-4, periodic=True, gravity=(0.0, 0.0, 0.0)))
    ref = {}
    for name, g in (("fast", fast), ("slow", slow)):
        _, b0 = g.neighbor_stats()
        g.step(100)
        _, b1 = g.neighbor_stats()
        ref[name] = b1 - b0
    assert ref["fast"] > ref["slow"]
    assert ref["fast"] >= 5


@pytest.mark.parametrize("precision", [0, MIXED])
@pytest.mark.parametrize("periodic", [False, True])
def test_step_torsor_equals_force_evaluation(gpu, precision, periodic):
    """The torsor the launch that ends a step leaves against a separate force
    evaluation at the same state; elastic frictionless contacts, so that the
    force does not depend on the half-step velocities the step used."""
    wl = workloads.dense_pack(side=8, en=1.0, muc=0.0, periodic=periodic,
                              gravity=(0.0, 0.0, 0.0) if periodic else (0.0, 0.0, -9.81))
    g = make_gpu(wl, precision=precision)
    assert g.neighbor_stats()[0] > 0
    for _ in range(3):
        g.step(40)
        s = g.state()
        g.compute_forces()
        f = g.state()
        assert np.array_equal(s["F"], f["F"]) and np.array_equal(s["T"], f["T"])
    assert np.abs(s["F"]).max() > 0
