"""The device narrow phase (RigidBodyWithCrust::ClosestPoint: sphere-sphere
and sphere-box closed forms, GJK on crust-shrunk bodies, ContactError
handling) compiled for the host through grains_gpu_closest_point, against
the oracle's restatement, for every pair of shape kinds: boxes, regular
tetrahedra, icosahedra and spheres at random orientations and separations
straddling contact.  Contact status, contact point, overlap vector and
overlap distance agree bit for bit."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from pacific_amd import GrainsGPU, shapes

KINDS = ["box", "tet", "ico", "sph"]


@pytest.fixture(scope="module")
def both():
    o = Oracle((0, 0, 0), (1, 1, 1), (0, 0, 0), 1.0, (0, 0, 0), 0.0, 1e-5, 0)
    g = GrainsGPU((0, 0, 0), (1, 1, 1))
    ids = {}
    for name in KINDS:
        if name == "box":
            a, b = o.add_shape(1, L=(2e-3, 2e-3, 2e-3)), g.add_shape(1, L=(2e-3, 2e-3, 2e-3))
        elif name == "sph":
            a, b = o.add_shape(0, radius=1e-3), g.add_shape(0, radius=1e-3)
        else:
            v, f = shapes.tetrahedron(1.2e-3) if name == "tet" else shapes.icosahedron(1.2e-3)
            a, b = o.add_shape(2, verts=v, faces=f), g.add_shape(2, verts=v, faces=f)
        assert a == b
        ids[name] = a
    return o, g, ids


@pytest.mark.parametrize("ka", KINDS)
@pytest.mark.parametrize("kb", KINDS)
def test_device_narrow_phase_equals_oracle(both, ka, kb):
    o, g, ids = both
    rng = np.random.default_rng(hash((ka, kb)) % 2**32)
    ncontact = 0
    for k in range(150):
        BA, BB = shapes.random_rotations(2, 1000 * KINDS.index(ka) + 10 * KINDS.index(kb) + k)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        xA = rng.uniform(-1e-3, 1e-3, 3)
        xB = xA + u * rng.uniform(1.0e-3, 3.2e-3)
        Po, ovo, do, so = o.closest_point(ids[ka], xA, BA, ids[kb], xB, BB)
        Pg, ovg, dg, sg = g.closest_point(ids[ka], xA, BA, ids[kb], xB, BB)
        assert so == sg, (k, so, sg)
        if so in (1, 2):
            assert np.array_equal(Po, Pg) and np.array_equal(ovo, ovg) and do == dg, k
            ncontact += do < 0
    assert ncontact > 20


@pytest.mark.parametrize("ka", KINDS)
@pytest.mark.parametrize("kb", KINDS)
def test_exchange_symmetry(both, ka, kb):
    """Each particle evaluates its pairs as body A, so a pair is evaluated
    from both sides on the device and once in the reference; the two
    evaluations must be mirror images: same point and distance, opposite
    overlap vector.  GJK is exactly symmetric under the exchange; the
    sphere-sphere closed form places the point from body A's centre, so its
    two evaluations agree to rounding."""
    _, g, ids = both
    rng = np.random.default_rng(7 + hash((ka, kb)) % 2**31)
    for k in range(100):
        BA, BB = shapes.random_rotations(2, 77 + 1000 * KINDS.index(ka) + 10 * KINDS.index(kb) + k)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        xA = rng.uniform(-1e-3, 1e-3, 3)
        xB = xA + u * rng.uniform(1.0e-3, 3.2e-3)
        P1, ov1, d1, s1 = g.closest_point(ids[ka], xA, BA, ids[kb], xB, BB)
        P2, ov2, d2, s2 = g.closest_point(ids[kb], xB, BB, ids[ka], xA, BA)
        assert s1 == s2, k
        if s1 in (1, 2):
            assert d1 == d2, (k, d1, d2)
            assert np.array_equal(ov1, -ov2), k
            if (ka, kb) == ("sph", "sph"):
                assert np.abs(P1 - P2).max() <= 4 * np.spacing(np.abs(P1).max()), k
            else:
                assert np.array_equal(P1, P2), (k, P1 - P2)


OPTIONS = [dict(gjk_sv=True), dict(tolerance=1e-6), dict(acceleration=True), dict(gjk_sv=True, acceleration=True),
           dict(bounding_volume=1), dict(bounding_volume=2), dict(gjk_sv=True, bounding_volume=1)]


@pytest.mark.parametrize("opts", OPTIONS, ids=lambda d: "-".join(f"{k}={v}" for k, v in d.items()))
def test_collision_options_equal_oracle(opts):
    """<CollisionDetectionAlgorithm> options (GJK_SV, Tolerance,
    Acceleration, BoundingVolume OBB / OBC) on both sides: the same bits."""
    o = Oracle((0, 0, 0), (1, 1, 1), (0, 0, 0), 1.0, (0, 0, 0), 0.0, 1e-5, 0)
    g = GrainsGPU((0, 0, 0), (1, 1, 1))
    ids = {}
    for name in KINDS:
        if name == "box":
            a, b = o.add_shape(1, L=(2e-3, 1.6e-3, 1.2e-3)), g.add_shape(1, L=(2e-3, 1.6e-3, 1.2e-3))
        elif name == "sph":
            a, b = o.add_shape(0, radius=1e-3), g.add_shape(0, radius=1e-3)
        else:
            v, f = shapes.tetrahedron(1.2e-3) if name == "tet" else shapes.icosahedron(1.2e-3)
            a, b = o.add_shape(2, verts=v, faces=f), g.add_shape(2, verts=v, faces=f)
        ids[name] = a
    o.collision(**opts)
    g.collision(**opts)
    rng = np.random.default_rng(11)
    status = {}
    for ka in KINDS:
        for kb in KINDS:
            for k in range(60):
                BA, BB = shapes.random_rotations(2, 5000 + 100 * KINDS.index(ka) + 10 * KINDS.index(kb) + k)
                u = rng.normal(size=3)
                u /= np.linalg.norm(u)
                xA = rng.uniform(-1e-3, 1e-3, 3)
                xB = xA + u * rng.uniform(1.0e-3, 3.2e-3)
                Po, ovo, do, so = o.closest_point(ids[ka], xA, BA, ids[kb], xB, BB)
                Pg, ovg, dg, sg = g.closest_point(ids[ka], xA, BA, ids[kb], xB, BB)
                assert so == sg, (ka, kb, k, so, sg)
                status[so] = status.get(so, 0) + 1
                if so in (1, 2):
                    assert np.array_equal(Po, Pg) and np.array_equal(ovo, ovg) and do == dg, (ka, kb, k)
    assert status.get(1, 0) > 150 and status.get(0, 0) > 100


@pytest.mark.parametrize("accel", [False, True])
def test_signed_volume_fixed_point_against_thin_wall(accel):
    """GJK_SV between a body and a long thin box (a silo wall 2 mm x 0.1 m x
    1.03 m) can settle on a step that re-adds and drops the same support
    point, which the loop repeats up to its cap of 1000 steps; the device
    ends such a loop as soon as a step leaves the simplex and v unchanged.
    Bodies near the wall at random poses: the oracle reaches the cap on some
    of them, and both sides give the same bits on all."""
    from pacific_amd import workloads
    wl = workloads.silo(200000)
    Lb, c, _, cr, _ = wl["boxes"][3]
    o = Oracle((0, 0, 0), (1, 1, 1), (0, 0, 0), 1.0, (0, 0, 0), 0.0, 1e-5, 0)
    g = GrainsGPU((0, 0, 0), (1, 1, 1))
    ids = []
    for s in wl["shapes"]:
        kw = dict(L=s.get("L"), verts=s.get("verts"), faces=s.get("faces"), crust=s["crust"])
        ids.append(o.add_shape(s["kind"], **kw))
        assert g.add_shape(s["kind"], **kw) == ids[-1]
    wall = o.add_shape(1, L=Lb, crust=cr)
    assert g.add_shape(1, L=Lb, crust=cr) == wall
    o.collision(gjk_sv=True, acceleration=accel)
    g.collision(gjk_sv=True, acceleration=accel)
    rng = np.random.default_rng(3)
    eye = np.eye(3).reshape(9)
    capped, status = 0, {}
    for k in range(2000):
        sh = k % len(ids)
        B = shapes.random_rotations(1, 900 + k)[0]
        x = np.array([c[0] + Lb[0] / 2 + rng.uniform(0.6e-3, 1.3e-3), c[1] + rng.uniform(-0.04, 0.04),
                      rng.uniform(0.001, 0.02)])
        Po, ovo, do, so = o.closest_point(ids[sh], x, B, wall, c, eye)
        capped += o.gjk_last_iterations() == 1000
        Pg, ovg, dg, sg = g.closest_point(ids[sh], x, B, wall, c, eye)
        assert so == sg, (k, so, sg)
        status[so] = status.get(so, 0) + 1
        assert np.array_equal(Po, Pg) and np.array_equal(ovo, ovg) and do == dg, k
    assert capped >= 20 and status.get(1, 0) > 100, (capped, status)
