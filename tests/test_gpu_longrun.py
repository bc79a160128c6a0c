"""Long-run statistics of the cooling granular gas (workloads.periodic_gas:
phi 0.3, r ~ U[0.8, 1.2] mm, sigma_v 0.1 m/s, Hooke kn 1e4 en 0.8 mu 0.3,
dt 1e-5, triply periodic, LeapFrog), 20,000 steps from the lattice start.

A many-body gas is chaotic: rounding differences (another contact-sum order,
FP32 contact arithmetic) grow until single trajectories decorrelate, so the
bars are on statistics.  Their size is set by the oracle itself: two oracle
runs of the 8,000-sphere gas whose initial states differ by 1e-12 m/s in one
velocity component agree in kinetic energy to 2.3e-13 at t = 0.01 s and
2.5e-10 at 0.02 s, and differ by up to 5.1e-3 afterwards.

* 8,000 spheres, device FP64 against the oracle's run (fixture
  tests/golden/gas8000_ke.json, written by tests/golden/make_gas8000_ke.py):
  kinetic energy every 1,000 steps within 1e-8 up to t = 0.02 s and within
  2e-2 (four times the oracle's own spread) over the whole run.
* 1,000,000 spheres, mixed precision against FP64: kinetic energy within
  5e-3 at every sample (measured 1.0e-3); Haff's law T0 / (1 + t/tau)^2
  fitted to each run after t = 0.05 s (KE^-1/2 linear in t) gives cooling
  times within 1% of each other (measured 0.15%), each fit within 10% of its
  samples (measured 7.1%: the gas drifts from homogeneous cooling as it
  clusters); the mean coordination number at t = 0.05, 0.1, 0.15, 0.2 s
  within 10% (measured up to 3.2%; a few thousand contacts per sample)."""
import json
import os

import numpy as np
import pytest

from pacific_amd import FP64, MIXED, workloads
from tests.helpers import make_gpu

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def kinetic_energy(g, m, I):
    s = g.state()
    return float(0.5 * np.sum(m[:, None] * s["v"] ** 2) + 0.5 * np.sum(I[:, None] * s["w"] ** 2))


def haff_fit(t, ke):
    """tau and the largest relative deviation of T0 / (1 + t/tau)^2 fitted
    through KE^-1/2 = a + b t (tau = a / b, T0 = a^-2)."""
    y = ke ** -0.5
    b, a = np.polyfit(t, y, 1)
    model = 1.0 / (a + b * t) ** 2
    return a / b, float(np.max(np.abs(model - ke) / ke))


def test_gas8000_kinetic_energy_against_oracle(gpu):
    ref = json.load(open(os.path.join(HERE, "golden", "gas8000_ke.json")))
    ke_ref = np.array(ref["kinetic_energy"])
    wl = workloads.periodic_gas(8000)
    g = make_gpu(wl, record=False)
    m, I, _ = g.masses()
    ke = [kinetic_energy(g, m, I)]
    for _ in range(20):
        g.step(ref["every_steps"])
        ke.append(kinetic_energy(g, m, I))
    rel = np.abs(np.array(ke) - ke_ref) / ke_ref
    print("8000-sphere gas, relative KE difference to the oracle:", " ".join(f"{r:.1e}" for r in rel))
    assert rel[:3].max() <= 1e-8, rel[:3]
    assert rel.max() <= 2e-2, rel


def test_cooling_gas_mixed_against_fp64_statistics(gpu):
    wl = workloads.periodic_gas(1_000_000)
    t = np.arange(21) * 1000 * wl["dt"]
    runs = {}
    for prec in (FP64, MIXED):
        g = make_gpu(wl, precision=prec, record=False)
        m, I, _ = g.masses()
        ke, coord = [kinetic_energy(g, m, I)], []
        for c in range(20):
            g.step(1000)
            ke.append(kinetic_energy(g, m, I))
            if (c + 1) % 5 == 0:
                coord.append(float(g.coordination_at_state().mean()))
        runs[prec] = (np.array(ke), np.array(coord))
        del g
    (k64, c64), (k32, c32) = runs[FP64], runs[MIXED]
    ke_rel = np.abs(k32 - k64) / k64
    late = t >= 0.05
    tau64, dev64 = haff_fit(t[late], k64[late])
    tau32, dev32 = haff_fit(t[late], k32[late])
    print(f"KE mixed vs FP64 max {ke_rel.max():.2e}; Haff tau {tau64:.5f} / {tau32:.5f} s, fit deviation "
          f"{dev64:.3f} / {dev32:.3f}; coordination {c64} / {c32}")
    assert ke_rel.max() <= 5e-3
    assert k64[-1] < 0.05 * k64[0]  # it cools
    assert abs(tau32 / tau64 - 1.0) <= 1e-2
    assert dev64 <= 0.1 and dev32 <= 0.1
    assert np.all(c64 > 0) and np.all(np.abs(c32 - c64) / c64 <= 0.1)
