"""Writes gas8000_ke.json: the kinetic energy of the FP64 oracle's run of an
8,000-sphere cooling gas (workloads.periodic_gas(8000): phi 0.3, r ~ U[0.8,
1.2] mm, sigma_v 0.1 m/s, Hooke kn 1e4 en 0.8 mu 0.3, dt 1e-5, triply
periodic) every 1,000 steps over 20,000 steps, for the long-run test of the
device against the oracle (tests/test_gpu_longrun.py).  Run from the
repository root: python3 tests/golden/make_gas8000_ke.py (about 2-3 minutes
on one core)."""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from pacific_amd import workloads  # noqa: E402
from tests.helpers import make_oracle  # noqa: E402


def kinetic_energy(o, m, I):
    s = o.state()
    return float(0.5 * np.sum(m[:, None] * s["v"] ** 2) + 0.5 * np.sum(I[:, None] * s["w"] ** 2))


def main():
    wl = workloads.periodic_gas(8000)
    o = make_oracle(wl)
    m, I, _ = o.masses()
    ke = [kinetic_energy(o, m, I)]
    for _ in range(20):
        o.step(1000)
        ke.append(kinetic_energy(o, m, I))
    out = {"workload": "periodic_gas(8000)", "every_steps": 1000, "dt": wl["dt"], "kinetic_energy": ke}
    with open(os.path.join(ROOT, "tests", "golden", "gas8000_ke.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
