"""Pins the CPU oracle (oracle/grains_oracle.cpp) against what the mathematics
fixes: exact trajectories of the integrators under constant loads, the
Hooke-dashpot-Coulomb closed forms (restitution, contact duration, static
overlap, Coulomb cap), conservation laws, and independent closest-point
geometry.  Each check is chosen so that a plausible slip (a dropped term, a
sign, a transposed operand, a wrong index) fails it.

Reference closed forms: HookeContactForceModel::computeAndWriteEstimates and
computeDeltaMax (HookeContactForceModel.cpp:208-277)."""
import math

import numpy as np
import pytest

from oracle.oracle import (ADAMS_BASHFORTH2, FIRST_ORDER_EXPLICIT, LEAPFROG, SECOND_ORDER_EXPLICIT, Oracle,
                           OracleError)

G = 9.81
RHO = 2500.0


def one_sphere(scheme=LEAPFROG, x=(0.01, 0.01, 0.05), v=(0, 0, 0), w=(0, 0, 0), dt=1e-5, gravity=(0, 0, -G),
               box=(0.02, 0.02, 0.06)):
    o = Oracle((0, 0, 0), box, gravity=gravity, dt=dt, scheme=scheme)
    o.contact_model(0, 0, 1e4, 0.9, 0.5)
    o.add_spheres(1e-3, RHO, 5e-5, 0, [x], [v], [w])
    o.init()
    return o


def test_mass_inertia_weight():
    o = one_sphere()
    m, I, wgt = o.masses()
    r = 1e-3
    assert m[0] == pytest.approx(RHO * 4.0 / 3.0 * math.pi * r ** 3, rel=1e-15)
    assert I[0] == pytest.approx(0.4 * m[0] * r * r, rel=1e-14)
    assert np.allclose(wgt[0], [0, 0, -G * m[0]], rtol=1e-15)


@pytest.mark.parametrize("n", [1, 7, 500])
def test_leapfrog_free_fall_is_exact_parabola(n):
    # kick-drift-kick is exact for constant acceleration: x_n = x0 + g t^2/2
    o = one_sphere()
    o.step(n)
    s = o.state()
    t = n * 1e-5
    assert s["x"][0, 2] == pytest.approx(0.05 - 0.5 * G * t * t, rel=1e-14, abs=1e-15)
    assert s["v"][0, 2] == pytest.approx(-G * t, rel=1e-12)
    assert s["x"][0, 0] == 0.01 and s["v"][0, 0] == 0.0


def test_first_order_explicit_free_fall():
    # x_{k+1} = x_k + v_k dt, v_{k+1} = v_k + g dt  =>  x_n = x0 + g dt^2 n(n-1)/2
    n, dt = 300, 1e-5
    o = one_sphere(FIRST_ORDER_EXPLICIT)
    o.step(n)
    s = o.state()
    assert s["x"][0, 2] == pytest.approx(0.05 - G * dt * dt * n * (n - 1) / 2, rel=1e-13)
    assert s["v"][0, 2] == pytest.approx(-G * n * dt, rel=1e-12)


def test_second_order_explicit_free_fall():
    n, dt = 300, 1e-5
    o = one_sphere(SECOND_ORDER_EXPLICIT)
    o.step(n)
    s = o.state()
    assert s["x"][0, 2] == pytest.approx(0.05 - 0.5 * G * (n * dt) ** 2, rel=1e-13)


def test_adams_bashforth_free_fall():
    # Euler start, then dx_k = (k - 1/2) g dt^2  =>  x_n = x0 + g dt^2 (n^2 - 1)/2
    n, dt = 300, 1e-5
    o = one_sphere(ADAMS_BASHFORTH2)
    o.step(n)
    s = o.state()
    assert s["x"][0, 2] == pytest.approx(0.05 - G * dt * dt * (n * n - 1) / 2, rel=1e-13)
    assert s["v"][0, 2] == pytest.approx(-G * n * dt, rel=1e-12)


def quat_axis_angle(axis, angle):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.concatenate([axis * math.sin(angle / 2), [math.cos(angle / 2)]])


def quat_mul(a, b):  # [x, y, z, w]
    av, aw, bv, bw = a[:3], a[3], b[:3], b[3]
    return np.concatenate([np.cross(av, bv) + aw * bv + bw * av, [aw * bw - av @ bv]])


def test_free_rotation_quaternion():
    # torque-free sphere: q(t) = rotation by |w| t about w; w stays w0
    w0 = np.array([30.0, -20.0, 45.0])
    o = one_sphere(w=w0, gravity=(0, 0, 0))
    n = 400
    o.step(n)
    s = o.state()
    q = s["q"][0]
    qref = quat_axis_angle(w0, np.linalg.norm(w0) * n * 1e-5)
    assert np.allclose(q, qref, rtol=0, atol=1e-13)
    assert abs(np.linalg.norm(q) - 1.0) < 1e-13  # rounding random walk over 400 products
    assert np.allclose(s["w"][0], w0, rtol=1e-13)


def test_ssprk3_constant_axial_torque():
    # torque along the spin axis: body-frame torque stays equal to it, so the
    # angular velocity grows linearly, w_n = w0 + n T/I dt, through KDK halves
    o = one_sphere(w=(0, 0, 10.0), gravity=(0, 0, 0))
    m, I, _ = o.masses()
    Tz = 3e-9
    dt = 1e-5
    for _ in range(50):
        o.set_forces(np.zeros((1, 3)), np.array([[0.0, 0.0, Tz]]))
        o.move_only(0.5 * dt, dt)
        o.advance_only(0.5 * dt)
    s = o.state()
    assert s["w"][0, 2] == pytest.approx(10.0 + 50 * Tz / I[0] * dt, rel=1e-13)
    assert abs(s["w"][0, 0]) < 1e-14 and abs(s["w"][0, 1]) < 1e-14


def beta(en):
    return math.log(en) / math.sqrt(math.pi ** 2 + math.log(en) ** 2)


@pytest.mark.parametrize("en", [0.9, 0.5])
def test_binary_collision_restitution_and_duration(en):
    # head-on collision of two equal spheres, no gravity: v_out/v_in = en,
    # duration tc = pi / sqrt(w0^2 - eta^2), w0 = sqrt(kn/m*), eta = -beta w0
    kn, r = 1e4, 1e-3
    m = RHO * 4 / 3 * math.pi * r ** 3
    mstar = m / 2
    w0 = math.sqrt(kn / mstar)
    eta = -beta(en) * w0
    tc = math.pi / math.sqrt(w0 ** 2 - eta ** 2)
    dt = tc / 2000
    o = Oracle((0, 0, 0), (0.02, 0.02, 0.02), gravity=(0, 0, 0), dt=dt)
    o.contact_model(0, 0, kn, en, 0.5)
    v0 = 0.2
    o.add_spheres(r, RHO, 5e-5, 0, [[0.009 - 1e-6, 0.01, 0.01], [0.011, 0.01, 0.01]],
                  [[v0 / 2, 0, 0], [-v0 / 2, 0, 0]])
    o.init()
    touching = []
    for k in range(4000):
        o.step(1)
        touching.append(len(o.pairs()) > 0)
    s = o.state()
    vrel = s["v"][1, 0] - s["v"][0, 0]
    assert vrel / v0 == pytest.approx(en, rel=2e-3)
    assert sum(touching) * dt == pytest.approx(tc, rel=2e-3)
    # momentum is conserved exactly up to rounding
    assert abs(s["v"][0, 0] + s["v"][1, 0]) < 1e-15


def test_sphere_wall_restitution():
    # sphere on a fixed wall: m* = m (wall mass 1e20)
    kn, en, r = 1e4, 0.7, 1e-3
    m = RHO * 4 / 3 * math.pi * r ** 3
    w0 = math.sqrt(kn / m)
    tc = math.pi / math.sqrt(w0 ** 2 - (beta(en) * w0) ** 2)
    dt = tc / 2000
    o = Oracle((0, 0, 0), (0.02, 0.02, 0.02), gravity=(0, 0, 0), dt=dt)
    o.contact_model(0, 0, kn, en, 0.5)
    o.add_spheres(r, RHO, 5e-5, 0, [[0.01, 0.01, r + 1e-6]], [[0, 0, -0.1]])
    o.add_box((0.021, 0.021, 0.0005), (0.01, 0.01, -0.00025))
    o.init()
    o.step(5000)
    s = o.state()
    assert s["v"][0, 2] / 0.1 == pytest.approx(en, rel=2e-3)


def test_static_overlap_on_floor():
    # a sphere at rest on the floor carries its weight: delta = m g / kn
    kn, r = 1e4, 1e-3
    o = Oracle((0, 0, 0), (0.02, 0.02, 0.02), dt=1e-5)
    o.contact_model(0, 0, kn, 0.5, 0.5)
    o.add_spheres(r, RHO, 5e-5, 0, [[0.01, 0.01, r + 1e-7]])
    o.add_box((0.021, 0.021, 0.0005), (0.01, 0.01, -0.00025))
    o.init()
    o.step(20000)
    s = o.state()
    m = RHO * 4 / 3 * math.pi * r ** 3
    delta = r - s["x"][0, 2]
    assert delta == pytest.approx(m * G / kn, rel=1e-6)
    assert abs(s["F"][0, 2]) < 1e-9 * m * G
    Fw, Tw = o.obstacle_loads()
    assert Fw[0, 2] == pytest.approx(-m * G, rel=1e-6)


def test_coulomb_cap_is_exact():
    # fast tangential sliding on the floor: |F_t| = mu |F_n| exactly, opposite
    # to the slip velocity
    kn, r, mu = 1e4, 1e-3, 0.3
    o = Oracle((0, 0, 0), (0.02, 0.02, 0.02), gravity=(0, 0, 0), dt=1e-5)
    o.contact_model(0, 0, kn, 0.8, mu)
    # viscous tangential force 2|beta| sqrt(kn m) |v_t| ~ 0.15 N > mu Fn = 0.06 N
    o.add_spheres(r, RHO, 5e-5, 0, [[0.01, 0.01, r - 2e-5]], [[3.0, 1.2, 0]])
    o.add_box((0.021, 0.021, 0.0005), (0.01, 0.01, -0.00025))
    o.init()
    s = o.state()
    F = s["F"][0]
    Fn, Ft = F[2], np.hypot(F[0], F[1])
    assert Fn == pytest.approx(kn * 2e-5, rel=1e-12)
    assert Ft == pytest.approx(mu * Fn, rel=1e-14)
    slip = np.array([3.0, 1.2])
    assert np.allclose(F[:2] / Ft, -slip / np.linalg.norm(slip), atol=1e-14)
    # the friction torque is r_contact x F_t with the lever pointing down
    lever = np.array([0, 0, -(r - 1e-5)])
    assert np.allclose(s["T"][0], np.cross(lever, [F[0], F[1], 0]) + 0.0, rtol=1e-6, atol=1e-18)


def random_packing(n, box, rmin, rmax, seed, overlap_frac=0.3):
    rng = np.random.default_rng(seed)
    r = rng.uniform(rmin, rmax, n)
    x = rng.uniform(rmax, np.asarray(box) - rmax, (n, 3))
    return r, x


def test_newton_third_law_and_angular_momentum():
    # dense random packing with small overlaps (no walls, no gravity): the
    # contact forces sum to zero and so do the moments about the origin
    rng = np.random.default_rng(3)
    m = 9
    a = 2.0e-3
    idx = np.stack(np.meshgrid(range(m), range(m), range(m), indexing="ij"), -1).reshape(-1, 3)
    x = (idx + 0.5) * a + rng.uniform(-2e-5, 2e-5, (len(idx), 3))
    r = rng.uniform(0.995e-3, 1.02e-3, len(idx))
    v = rng.normal(0, 0.05, (len(idx), 3))
    w = rng.normal(0, 20, (len(idx), 3))
    o = Oracle((0, 0, 0), (m * a, m * a, m * a), gravity=(0, 0, 0), dt=1e-5)
    o.contact_model(0, 0, 1e4, 0.8, 0.4, kr=0.02)
    o.add_spheres(r, RHO, 5e-5, 0, x, v, w)
    o.init()
    s = o.state()
    assert len(o.pairs()) > 1000
    scale = np.sum(np.abs(s["F"]))
    assert np.max(np.abs(s["F"].sum(axis=0))) < 1e-13 * scale
    L = s["T"] + np.cross(s["x"], s["F"])
    lscale = np.sum(np.abs(np.cross(s["x"], s["F"])))
    assert np.max(np.abs(L.sum(axis=0))) < 1e-13 * lscale


@pytest.mark.parametrize("w1", [0.0, 40.0])
def test_rolling_resistance_moment(w1):
    # HookeContactForceModel.cpp:95-111: M = -(kr Req |F_n| |w0 x (c0-P) -
    # w1 x (c1-P)| / |w0 - w1|) (w0 - w1) on body 0, -M on body 1.  Two equal
    # spheres meet head on along x, no gravity, frictionless (muc = 0), spun
    # about z.  The normal force passes through both centres, so M is the only
    # torque.  With the contact point half way (lever l = r - delta/2 on both
    # sides) |w0 x (c0-P) - w1 x (c1-P)| = l |w0 + w1|, so w0 + w1 is conserved
    # and d(w0 - w1)/dt = -2 kr Req l |F_n| (w0 + w1) / I: over the collision
    # the spin difference drops by 2 kr Req r J (w0 + w1) / I up to the lever's
    # relative excursion delta_max / 2r (~2e-3 here), with J the normal
    # impulse, Req = r/2 and I = 2/5 m r^2.  A wrong Req, a dropped lever
    # factor or a sign slip each miss by a factor.
    kn, en, r, kr = 1e4, 0.9, 1e-3, 0.6
    m = RHO * 4 / 3 * math.pi * r ** 3
    inertia = 0.4 * m * r ** 2
    mstar = m / 2
    w0n = math.sqrt(kn / mstar)
    tc = math.pi / math.sqrt(w0n ** 2 - (beta(en) * w0n) ** 2)
    dt = tc / 2000
    o = Oracle((0, 0, 0), (0.02, 0.02, 0.02), gravity=(0, 0, 0), dt=dt)
    o.contact_model(0, 0, kn, en, 0.0, kr=kr)
    u = 0.2
    w0 = 100.0
    o.add_spheres(r, RHO, 5e-5, 0, [[0.009 - 1e-6, 0.01, 0.01], [0.011, 0.01, 0.01]],
                  [[u / 2, 0, 0], [-u / 2, 0, 0]], [[0, 0, w0], [0, 0, w1]])
    o.init()
    o.step(4000)
    s = o.state()
    assert len(o.pairs()) == 0
    J = m * (u / 2 - s["v"][0, 0])  # impulse on body 0 (its velocity drops by J/m)
    assert J == pytest.approx(mstar * (1 + en) * u, rel=3e-3)
    drop = (w0 - w1) - (s["w"][0, 2] - s["w"][1, 2])
    expected = 2 * kr * (r / 2) * r * J * (w0 + w1) / inertia
    assert 0.2 < drop / (w0 - w1) < 0.9  # a sizeable, partial decay
    assert drop == pytest.approx(expected, rel=3e-3)
    # w0 + w1 is conserved; no spin appears about x or y; no tangential force
    assert s["w"][0, 2] + s["w"][1, 2] == pytest.approx(w0 + w1, rel=1e-12)
    assert np.max(np.abs(s["w"][:, :2])) == 0.0
    assert np.max(np.abs(s["v"][:, 1:])) == 0.0


def closest_point_box(c, L, centre, B):
    """independent sphere-box geometry: clamp the centre into the box frame"""
    B = np.asarray(B).reshape(3, 3)
    local = B.T @ (np.asarray(c) - centre)
    e = np.asarray(L) / 2
    q = np.clip(local, -e, e)
    return centre + B @ q


@pytest.mark.parametrize("seed", range(4))
def test_sphere_box_geometry(seed):
    # overlap depth and normal against the clamp construction, for a rotated
    # box and sphere centres in face, edge and corner regions.  Overlaps stay
    # below the crust sum, the regime the reference admits (deeper ones are a
    # ContactError); there the edge/corner contact point of Box.cpp:725-769
    # stays between the sphere centre and the box.
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, 2 * np.pi, 3)
    cx, cy, cz = np.cos(ang)
    sx, sy, sz = np.sin(ang)
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    B = Rz @ Ry @ Rx
    L = np.array([4e-3, 3e-3, 2e-3])
    e = L / 2
    centre = np.array([0.01, 0.01, 0.01])
    kn, r, crust = 1e4, 1e-3, 5e-5
    # sample local centres in the shell r - 9e-5 < dist < r around the box
    loc = rng.uniform(-(e + r), e + r, (200000, 3))
    dist = np.linalg.norm(loc - np.clip(loc, -e, e), axis=1)
    loc = loc[(dist > r - 9e-5) & (dist < r - 1e-7)]
    outside = (np.abs(loc) > e).sum(axis=1)  # 1 face, 2 edge, 3 corner region
    picks = []
    for kind in (1, 2, 3):
        picks += list(loc[outside == kind][:4])
    assert len(picks) == 12
    for local in picks:
        c = centre + B @ local
        cp = closest_point_box(c, L, centre, B)
        dist = np.linalg.norm(c - cp)
        o = Oracle((0, 0, 0), (0.02, 0.02, 0.02), gravity=(0, 0, 0), dt=1e-5)
        o.contact_model(0, 0, kn, 0.8, 0.4)
        o.add_spheres(r, RHO, crust, 0, [c])
        o.add_box(L, centre, B.reshape(9), crust, 0)
        o.init()
        F = o.state()["F"][0]
        n = (c - cp) / dist
        assert np.linalg.norm(F) == pytest.approx(kn * (r - dist), rel=1e-8)
        assert np.allclose(F / np.linalg.norm(F), n, atol=1e-9)


def test_sphere_sphere_geometry():
    rng = np.random.default_rng(11)
    kn = 1e4
    for _ in range(20):
        ra, rb = rng.uniform(0.8e-3, 1.2e-3, 2)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        delta = rng.uniform(1e-6, 4e-5)
        xa = np.array([0.01, 0.01, 0.01])
        xb = xa + d * (ra + rb - delta)
        o = Oracle((0, 0, 0), (0.02, 0.02, 0.02), gravity=(0, 0, 0), dt=1e-5)
        o.contact_model(0, 0, kn, 0.8, 0.4)
        o.add_spheres([ra, rb], RHO, 5e-5, 0, [xa, xb])
        o.init()
        s = o.state()
        assert np.allclose(s["F"][0], -kn * delta * d, rtol=1e-9, atol=1e-12 * kn * delta)
        assert np.allclose(s["F"][1], kn * delta * d, rtol=1e-9, atol=1e-12 * kn * delta)
        assert np.allclose(s["T"], 0, atol=1e-15 * kn * delta)


def test_errors():
    # MotionError: displacement over one step larger than the crust
    o = one_sphere(v=(0, 0, -10.0))
    with pytest.raises(OracleError) as e:
        o.step(1)
    assert e.value.kind == "MotionError"
    # SimulationError: a particle leaves the grid
    o = one_sphere(x=(0.01, 0.01, 3e-5), v=(0, 0, -4.0))
    with pytest.raises(OracleError) as e:
        o.step(2)
    assert e.value.kind == "SimulationError"
    # ContactError: overlap beyond the crusts and beyond the recovery
    o = Oracle((0, 0, 0), (0.02, 0.02, 0.02), gravity=(0, 0, 0))
    o.contact_model(0, 0, 1e4, 0.8, 0.4)
    o.add_spheres(1e-3, RHO, 5e-5, 0, [[0.01, 0.01, 0.01], [0.0103, 0.01, 0.01]])
    with pytest.raises(OracleError) as e:
        o.init()
    assert e.value.kind == "ContactError"
    # overlap beyond the crusts but recoverable: imposed overlap = crust sum
    o = Oracle((0, 0, 0), (0.02, 0.02, 0.02), gravity=(0, 0, 0))
    o.contact_model(0, 0, 1e4, 0.8, 0.4)
    o.add_spheres(1e-3, RHO, 5e-5, 0, [[0.01, 0.01, 0.01], [0.0118, 0.01, 0.01]])
    o.init()
    s = o.state()
    assert o.stats()["recovered"] == 1
    assert np.linalg.norm(s["F"][0]) == pytest.approx(1e4 * 1e-4, rel=1e-12)
