"""The oracle's HookeMemory contact model (HookeMemoryContactForceModel.cpp)
pinned by rigid-body mechanics, without a device.

* a sphere released on a plane inclined at theta, with a tangential spring
  and friction above 2/7 tan(theta), rolls without slipping: its centre
  accelerates at 5/7 g sin(theta) and v = omega r;
* with rolling friction mur > tan(theta) the rolling spring holds the sphere
  at rest, the contact force cancelling the weight;
* the memoryless Hooke-dashpot-Coulomb law, whose tangential force vanishes
  with the slip velocity, lets a cube below its friction angle creep down at
  m g sin(theta) / gamma_t.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from pacific_amd import workloads

RHO, G = 2500.0, 9.81


def slope_system(theta, memory=True, muc=0.5):
    R = workloads._rot_y(theta)
    o = Oracle((-0.03, -0.03, -0.03), (0.03, 0.03, 0.03), gravity=(0, 0, -G), dt=2e-6)
    if memory:
        o.memory_contact_model(0, 0, 1e4, 0.3, muc, kt=2.0 / 7.0 * 1e4)
    else:
        o.contact_model(0, 0, 1e4, 0.3, muc)
    o.add_box((0.05, 0.05, 0.002), (0.0, 0.0, 0.0), R.ravel(), 5e-5, 0)
    return o, R


def down_slope(R):
    # unit vector along the plane pointing downhill
    e1 = R[:, 0]
    return e1 if e1[2] < 0 else -e1


@pytest.mark.parametrize("theta", [15.0, 25.0])
def test_sphere_rolls_without_slipping(theta):
    r = 1e-3
    o, R = slope_system(theta)
    n = R[:, 2]
    m = RHO * 4 / 3 * math.pi * r ** 3
    delta = m * G * math.cos(math.radians(theta)) / 1e4
    o.add_spheres(r, RHO, 5e-5, 0, [(0.001 + r - delta) * n])
    o.init()
    t_s, v_s = [], []
    d = down_slope(R)
    for k in range(40):
        o.step(500)
        s = o.state()
        t_s.append((k + 1) * 500 * 2e-6)
        v_s.append(float(s["v"][0] @ d))
    t_s, v_s = np.array(t_s[10:]), np.array(v_s[10:])
    a = np.polyfit(t_s, v_s, 1)[0]
    assert a == pytest.approx(5.0 / 7.0 * G * math.sin(math.radians(theta)), rel=1e-2)
    s = o.state()
    w_axis = np.cross(n, d)  # rolling downhill turns about n x d
    assert float(s["w"][0] @ w_axis) * r == pytest.approx(float(s["v"][0] @ d), rel=1e-2)


def test_rolling_spring_holds_sphere_on_slope():
    """Rolling friction with memory (mur > tan(theta)): the spring torque
    holds the sphere at rest, the contact force balancing the weight."""
    theta, r = 15.0, 1e-3
    R = workloads._rot_y(theta)
    o = Oracle((-0.03, -0.03, -0.03), (0.03, 0.03, 0.03), gravity=(0, 0, -G), dt=2e-6)
    o.memory_contact_model(0, 0, 1e4, 0.3, 0.5, kt=2.0 / 7.0 * 1e4, mur=0.5, etarpf=1.0)
    o.add_box((0.05, 0.05, 0.002), (0.0, 0.0, 0.0), R.ravel(), 5e-5, 0)
    n = R[:, 2]
    m = RHO * 4 / 3 * math.pi * r ** 3
    delta = m * G * math.cos(math.radians(theta)) / 1e4
    o.add_spheres(r, RHO, 5e-5, 0, [(0.001 + r - delta) * n])
    o.init()
    o.step(20000)
    s = o.state()
    d = down_slope(R)
    assert abs(float(s["v"][0] @ d)) < 1e-4
    assert np.linalg.norm(s["w"][0]) * r < 1e-4
    _, _, wgt = o.masses()
    assert np.abs(s["F"][0]).max() < 1e-3 * m * G  # contact force cancels the weight


def test_memoryless_cube_creeps_at_viscous_terminal_velocity():
    """The Hooke-dashpot-Coulomb tangential force vanishes with the slip
    velocity: a cube below its friction angle slides at m g sin(theta) /
    gamma_t, gamma_t = gamma_n when etat = -1."""
    theta, L = 20.0, 2e-3
    o, R = slope_system(theta, memory=False)
    cube = o.add_shape(1, L=(L, L, L))
    n = R[:, 2]
    o.add_particles(cube, RHO, 0, [(0.001 + L / 2 - 1e-6) * n], basis=[R.ravel()])
    o.init()
    o.step(20000)
    m = RHO * L ** 3
    v = float(o.state()["v"][0] @ down_slope(R))
    en = 0.3
    beta = math.log(en) / math.sqrt(math.pi ** 2 + math.log(en) ** 2)
    gammat = 2.0 * (-beta * math.sqrt(1e4 / m)) * m
    assert v == pytest.approx(m * G * math.sin(math.radians(theta)) / gammat, rel=5e-2)
