"""Non-spherical particles on the HIP path against the oracle: free rigid
rotation (3D kinematics, basis composition), contact lists and forces of a
settled mix of cubes, tetrahedra, icosahedra and spheres through GJK, and a
short rollout.

Bars: positions and velocities of free bodies bit-exact; quaternions and
bases within 1e-14 (device sin/cos); contact pair sets equal; forces and
torques within 1e-12 relative.  The narrow phase itself is bit-exact against
the oracle (tests/test_convex_narrow_phase.py); GJK(A, B) and GJK(B, A) give
the same bits (the iteration is symmetric under the exchange), so evaluating
each pair from both sides leaves only the order of a particle's force sum.
"""
import numpy as np
import pytest

from pacific_amd import workloads
from tests.helpers import contact_scale, make_gpu, make_oracle, pair_set, rel_force_error

pytestmark = pytest.mark.gpu


def transfer(o, g):
    so = o.state()
    g.set_basis(o.basis().reshape(-1, 9))
    g.set_state(x=so["x"], v=so["v"], w=so["w"], q=so["q"], recompute=True)
    o.set_state(x=so["x"], v=so["v"], w=so["w"], q=so["q"], recompute=True)


def test_free_rotation_bit_parity(gpu):
    wl = workloads.polyhedra_box((3, 3, 2), seed=5, sigma=0.2)
    wl["gravity"] = (0.0, 0.0, 0.0)
    wl["boxes"] = []
    # spread the bodies so that none touches another
    for gi, g in enumerate(wl["groups"]):
        x = g[1] * 3.0
        wl["groups"][gi] = (g[0], x) + tuple(g[2:])
    wl["max"] = (0.05, 0.05, 0.05)
    o, g = make_oracle(wl), make_gpu(wl)
    for _ in range(4):
        o.step(50)
        g.step(50)
        so, sg = o.state(), g.state()
        assert len(o.pairs()) == 0
        for k in ("x", "v"):
            assert np.array_equal(sg[k], so[k]), k
        assert np.abs(sg["q"] - so["q"]).max() < 1e-14
        assert np.abs(sg["w"] - so["w"]).max() <= 1e-13 * np.abs(so["w"]).max()
        assert np.abs(g.basis() - o.basis()).max() < 1e-14


@pytest.mark.parametrize("spheres", [False, True])
def test_settled_polyhedra_pairs_and_forces(gpu, spheres):
    wl = workloads.polyhedra_box((5, 5, 4), seed=9, spheres=spheres, sigma=0.2, en=0.3)
    o = make_oracle(wl)
    o.step(20000)
    g = make_gpu(wl)
    transfer(o, g)
    po, pg = pair_set(o.pairs()), pair_set(g.pairs())
    assert len(po) > 100
    assert pg == po
    so, sg = o.state(), g.state()
    _, _, wgt = o.masses()
    scale = contact_scale(so, wgt)
    tol = 1e-12
    eF = rel_force_error(sg["F"], so["F"], scale)
    eT = float(np.max(np.linalg.norm(sg["T"] - so["T"], axis=1) / (scale * 1.8e-3)))
    assert eF <= tol, eF
    assert eT <= tol, eT


def test_polyhedra_short_rollout(gpu):
    wl = workloads.polyhedra_box((4, 4, 3), seed=13, sigma=0.2, en=0.3)
    o = make_oracle(wl)
    o.step(12000)
    g = make_gpu(wl)
    transfer(o, g)
    o.step(300)
    g.step(300)
    so, sg = o.state(), g.state()
    assert np.abs(sg["x"] - so["x"]).max() < 1e-9
    assert np.abs(sg["v"] - so["v"]).max() < 1e-6 * max(np.abs(so["v"]).max(), 1e-3)


@pytest.mark.parametrize("opts", [dict(gjk_sv=True), dict(bounding_volume=1),
                                  dict(bounding_volume=2, acceleration=True), dict(gjk_sv=True, tolerance=1e-8)],
                         ids=["gjk_sv", "obb", "obc_accel", "gjk_sv_tol"])
def test_collision_options_pairs_and_forces(gpu, opts):
    """<CollisionDetectionAlgorithm> options on a settling packing of
    cubes, tetrahedra and icosahedra: the oracle's pair set and forces
    within 1e-12 after 12,000 oracle steps (the narrow phase itself is
    bit-identical, tests/test_convex_narrow_phase.py), and a short rollout
    within the bars of test_polyhedra_short_rollout."""
    wl = workloads.polyhedra_box((4, 4, 3), seed=13, sigma=0.2, en=0.3)
    o = make_oracle(wl, collision=opts)
    o.step(12000)
    g = make_gpu(wl, collision=opts)
    transfer(o, g)
    po, pg = pair_set(o.pairs()), pair_set(g.pairs())
    assert len(po) > 40
    assert pg == po
    so, sg = o.state(), g.state()
    _, _, wgt = o.masses()
    scale = contact_scale(so, wgt)
    assert rel_force_error(sg["F"], so["F"], scale) <= 1e-12
    o.step(300)
    g.step(300)
    so, sg = o.state(), g.state()
    assert np.abs(sg["x"] - so["x"]).max() < 1e-9
