"""The oracle's contact statistics (AppCollision::addToContactsFeatures,
AppCollision.cpp:159-189) pinned by closed forms, without a device.

* head-on collision of two equal spheres (linear damped oscillator in the
  overlap): maximum overlap (v0/wd) exp(-eta t*) sin(wd t*) at
  t* = atan2(wd, eta)/wd after first contact, mean overlap over the contact
  v0 (1 + exp(-eta tc)) / ((eta^2 + wd^2) tc), one overlap counted per step
  in contact, closed-form contact points counting one iteration;
* a pair across a periodic boundary is met twice per evaluation (master
  against clone from both sides), the same pair inside the box once.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle

RHO = 2500.0


def beta(en):
    return math.log(en) / math.sqrt(math.pi ** 2 + math.log(en) ** 2)


@pytest.mark.parametrize("en", [0.9, 0.5])
def test_binary_collision_overlap_statistics(en):
    kn, r = 1e4, 1e-3
    m = RHO * 4 / 3 * math.pi * r ** 3
    w0 = math.sqrt(kn / (m / 2))
    eta = -beta(en) * w0
    wd = math.sqrt(w0 ** 2 - eta ** 2)
    tc = math.pi / wd
    dt = tc / 4000
    v0, gap = 0.2, 1e-6
    o = Oracle((0, 0, 0), (0.02, 0.02, 0.02), gravity=(0, 0, 0), dt=dt)
    o.contact_model(0, 0, kn, en, 0.5)
    o.add_spheres(r, RHO, 5e-5, 0, [[0.009 - gap, 0.01, 0.01], [0.011, 0.01, 0.01]],
                  [[v0 / 2, 0, 0], [-v0 / 2, 0, 0]])
    o.init()
    touching = 0
    for _ in range(6000):
        o.step(1)
        touching += len(o.pairs()) > 0
    st = o.stats()
    tstar = math.atan2(wd, eta) / wd
    dmax = v0 / wd * math.exp(-eta * tstar) * math.sin(wd * tstar)
    dmean = v0 * (1 + math.exp(-eta * tc)) / ((eta ** 2 + wd ** 2) * tc)
    assert st["overlap_max"] == pytest.approx(dmax, rel=1e-3)
    assert st["time_overlap_max"] == pytest.approx(gap / v0 + tstar, abs=3 * dt)
    assert st["overlap_mean"] == pytest.approx(dmean, rel=2e-3)
    assert st["n_overlaps"] == touching
    assert st["gjk_iter_mean"] == 1.0 and st["n_evaluations"] == touching


@pytest.mark.parametrize("across", [True, False])
def test_periodic_pair_counted_twice(across):
    r, L, ov = 1e-3, 0.02, 2e-5
    o = Oracle((0, 0, 0), (L, L, L), (1, 0, 0), 1.0, (0, 0, 0), 0.0, 1e-6, 0)
    o.contact_model(0, 0, 1e4, 0.8, 0.5)
    xa = 0.5 * r if across else 0.01
    xb = xa - (2 * r - ov)
    if across:
        xb += L
    o.add_spheres(r, RHO, 5e-5, 0, [[xa, 0.01, 0.01], [xb, 0.01, 0.01]])
    o.init()
    st = o.stats()
    assert st["n_overlaps"] == (2 if across else 1)
    assert st["overlap_max"] == pytest.approx(ov, rel=1e-9)
    assert st["time_overlap_max"] == 0.0
