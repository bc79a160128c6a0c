"""Derivative code of one Hooke contact (host copy of the adjoint's
vector-Jacobian products, ``grains_contact_eval``), checked without a device:

* its forward loads equal the oracle's two-body forces and torques
  (sphere-sphere; sphere-box faces, edges and corners, rotated box), which
  pins the exact-arithmetic restatement the derivative is taken of;
* its gradients equal central finite differences of those loads with respect
  to body 0's position, velocity and angular velocity and to kn, en, muc, in
  the sticking and the sliding (Coulomb) regimes, with en-derived and given
  tangential damping.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from pacific_amd.grains import contact_eval

RHO, CRUST = 2500.0, 5e-5


def sphere_mass(r):
    return 4.0 / 3.0 * np.pi * r ** 3 * RHO


def two_sphere_oracle(x0, r0, v0, w0, x1, r1, v1, w1, model):
    kn, en, etat, muc = model
    o = Oracle((-0.01, -0.01, -0.01), (0.01, 0.01, 0.01), (0, 0, 0), 1.0, (0.0, 0.0, 0.0), 0.0, 1e-6, 0)
    o.contact_model(0, 0, kn, en, muc, etat=etat)
    o.add_spheres(np.array([r0, r1]), np.full(2, RHO), np.full(2, CRUST), np.zeros(2, np.int32),
                  np.array([x0, x1]), np.array([v0, v1]), np.array([w0, w1]))
    o.init()
    return o.state()


def sphere_box_oracle(x0, r0, v0, w0, L, c, B, model):
    kn, en, etat, muc = model
    o = Oracle((-0.01, -0.01, -0.01), (0.01, 0.01, 0.01), (0, 0, 0), 1.0, (0.0, 0.0, 0.0), 0.0, 1e-6, 0)
    o.contact_model(0, 0, kn, en, muc, etat=etat)
    o.add_spheres(np.array([r0]), np.array([RHO]), np.array([CRUST]), np.zeros(1, np.int32), np.array([x0]),
                  np.array([v0]), np.array([w0]))
    o.add_box(np.array(L), np.array(c), np.array(B).ravel(), CRUST, 0)
    o.init()
    return o.state()


def rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def tangent(rng, u):
    t = np.cross(u, rng.normal(size=3))
    return t / np.linalg.norm(t)


def pair_case(rng, slide):
    """Two overlapping spheres; slide: a 4 m/s tangential relative velocity
    and a small overlap, so that the Coulomb cap is active."""
    r0, r1 = rng.uniform(0.8e-3, 1.2e-3, 2)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    ov = rng.uniform(2e-6, 8e-6) if slide else rng.uniform(2e-6, 4e-5)
    x0 = rng.uniform(-1e-3, 1e-3, 3)
    x1 = x0 + (r0 + r1 - ov) * u
    v0, w0, w1 = rng.normal(0, 1e-3, 3), rng.normal(0, 0.3, 3), rng.normal(0, 0.3, 3)
    v1 = v0 + rng.normal(0, 1e-3, 3) + (4.0 * tangent(rng, u) if slide else 0.0)
    return x0, r0, v0, w0, x1, r1, v1, w1


def sliding(loads, u, mu):
    """|FT| = mu |FN| on body 0 (Coulomb cap active)"""
    f = loads[:3]
    fn = abs(f @ u)
    ft = np.linalg.norm(f - (f @ u) * u)
    return abs(ft - mu * fn) < 1e-9 * fn


def body(x, r, v, w):
    return np.concatenate([x, [r], v, w, [sphere_mass(r)]])


MODELS = [(1e4, 0.8, -1.0, 0.3), (5e3, 0.6, -1.0, 0.7), (1e4, 0.9, 5e3, 0.4)]


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("slide", [False, True])
def test_pair_loads_match_oracle(seed, slide):
    rng = np.random.default_rng(100 + seed)
    model = MODELS[seed % len(MODELS)]
    x0, r0, v0, w0, x1, r1, v1, w1 = pair_case(rng, slide)
    st = two_sphere_oracle(x0, r0, v0, w0, x1, r1, v1, w1, model)
    hit, loads, _ = contact_eval(0, body(x0, r0, v0, w0), body(x1, r1, v1, w1), model)
    assert hit
    scaleF = np.linalg.norm(st["F"][0])
    scaleT = max(np.linalg.norm(st["T"][0]), scaleF * 1e-3)
    assert np.linalg.norm(loads[0:3] - st["F"][0]) <= 1e-12 * scaleF
    assert np.linalg.norm(loads[6:9] - st["F"][1]) <= 1e-12 * scaleF
    assert np.linalg.norm(loads[3:6] - st["T"][0]) <= 1e-11 * scaleT
    assert np.linalg.norm(loads[9:12] - st["T"][1]) <= 1e-11 * scaleT


def box_case(rng, feature, slide):
    """Sphere next to a rotated box, closest to a face (1 axis outside the
    extents), an edge (2) or a corner (3); returns also the unit vector u from
    the centre towards the feature."""
    L = rng.uniform(2e-3, 6e-3, 3)
    c = rng.uniform(-1e-3, 1e-3, 3)
    B = rotation(rng)
    r0 = rng.uniform(0.8e-3, 1.2e-3)
    e = L / 2
    axes = rng.permutation(3)[:feature]
    local = rng.uniform(-0.8, 0.8, 3) * e
    sgn = rng.choice([-1.0, 1.0], 3)
    dirn = np.zeros(3)
    for a in axes:
        local[a] = sgn[a] * e[a]
        dirn[a] = sgn[a] * rng.uniform(0.5, 1.0)
    dirn /= np.linalg.norm(dirn)
    ov = rng.uniform(2e-6, 8e-6) if slide else rng.uniform(2e-6, 4e-5)
    local = local + (r0 - ov) * dirn
    x0 = B @ local + c
    u = -(B @ dirn)
    v0 = rng.normal(0, 1e-3, 3) + (4.0 * tangent(rng, u) if slide else 0.0)
    return x0, r0, v0, rng.normal(0, 0.3, 3), L, c, B, u


@pytest.mark.parametrize("feature", [1, 2, 3])
@pytest.mark.parametrize("seed", range(3))
def test_box_loads_match_oracle(feature, seed):
    rng = np.random.default_rng(200 + 10 * feature + seed)
    model = MODELS[seed % len(MODELS)]
    x0, r0, v0, w0, L, c, B, _ = box_case(rng, feature, slide=seed % 2 == 1)
    st = sphere_box_oracle(x0, r0, v0, w0, L, c, B, model)
    hit, loads, _ = contact_eval(1, body(x0, r0, v0, w0), np.concatenate([L, c, B.ravel()]), model)
    assert hit
    scaleF = np.linalg.norm(st["F"][0])
    scaleT = max(np.linalg.norm(st["T"][0]), scaleF * 1e-3)
    assert np.linalg.norm(loads[0:3] - st["F"][0]) <= 1e-11 * scaleF
    assert np.linalg.norm(loads[3:6] - st["T"][0]) <= 1e-10 * scaleT


def fd_grad(kind, b0, b1, model, cot):
    """Central differences of cot . loads over body 0's x, v, w and kn, en, muc."""
    steps = [1e-10] * 3 + [1e-7] * 3 + [1e-4] * 3
    out, noise = np.zeros(12), np.zeros(12)

    def L(b, m):
        hit, loads, _ = contact_eval(kind, b, b1, m)
        assert hit
        return float(cot @ loads)

    slots = [0, 1, 2, 4, 5, 6, 7, 8, 9]
    for i, (sl, h) in enumerate(zip(slots, steps)):
        bp, bm = b0.copy(), b0.copy()
        bp[sl] += h
        bm[sl] -= h
        out[i] = (L(bp, model) - L(bm, model)) / (2 * h)
        noise[i] = 1e-13 * max(abs(cot) @ abs(contact_eval(kind, b0, b1, model)[1]), 1e-300) / h
    for i, sl in enumerate([0, 1, 3]):
        h = abs(model[sl]) * 1e-6
        mp, mm = list(model), list(model)
        mp[sl] += h
        mm[sl] -= h
        out[9 + i] = (L(b0, mp) - L(b0, mm)) / (2 * h)
        noise[9 + i] = 1e-13 * max(abs(cot) @ abs(contact_eval(kind, b0, b1, model)[1]), 1e-300) / h
    return out, noise


def check_grad(g, fd):
    # per block (x, v, w, each parameter): relative to the block's magnitude,
    # plus the rounding noise of the differences (1e-13 |L| / h)
    gfd, noise = fd
    for sl in (slice(0, 3), slice(3, 6), slice(6, 9), slice(9, 10), slice(10, 11), slice(11, 12)):
        scale = np.linalg.norm(gfd[sl])
        tol = 2e-5 * scale + np.linalg.norm(noise[sl])
        assert np.linalg.norm(g[sl] - gfd[sl]) <= tol, (sl, g[sl], gfd[sl])


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("slide", [False, True])
def test_pair_vjp_matches_finite_differences(seed, slide):
    rng = np.random.default_rng(300 + seed)
    model = MODELS[seed % len(MODELS)]
    x0, r0, v0, w0, x1, r1, v1, w1 = pair_case(rng, slide)
    b0, b1 = body(x0, r0, v0, w0), body(x1, r1, v1, w1)
    cot = rng.normal(size=12)
    cot[3:6] *= 1e3
    cot[9:12] *= 1e3
    hit, loads, g = contact_eval(0, b0, b1, model, cot)
    assert hit
    assert sliding(loads, (x1 - x0) / np.linalg.norm(x1 - x0), model[3]) == slide
    check_grad(g, fd_grad(0, b0, b1, model, cot))


@pytest.mark.parametrize("feature", [1, 2, 3])
@pytest.mark.parametrize("slide", [False, True])
def test_wall_vjp_matches_finite_differences(feature, slide):
    rng = np.random.default_rng(400 + 10 * feature + int(slide))
    model = MODELS[feature % len(MODELS)]
    x0, r0, v0, w0, L, c, B, u = box_case(rng, feature, slide)
    b0, b1 = body(x0, r0, v0, w0), np.concatenate([L, c, B.ravel()])
    cot = rng.normal(size=12)
    cot[3:6] *= 1e3
    cot[6:] = 0.0
    hit, loads, g = contact_eval(1, b0, b1, model, cot)
    assert hit
    assert sliding(loads, u, model[3]) == slide
    check_grad(g, fd_grad(1, b0, b1, model, cot))
