"""The committed kinetic-energy curve of the 8,000-sphere cooling gas
(tests/golden/gas8000_ke.json, the oracle's run, read by the long-run GPU
test) is the current oracle's: its first two samples are reproduced bit for
bit (the run is deterministic; later samples need the whole 20,000 steps,
about two minutes on one core, which tests/golden/make_gas8000_ke.py
regenerates)."""
import json
import os

import numpy as np

from pacific_amd import workloads
from tests.helpers import make_oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def test_gas8000_fixture_is_the_oracle_run():
    ref = json.load(open(os.path.join(HERE, "golden", "gas8000_ke.json")))
    wl = workloads.periodic_gas(8000)
    o = make_oracle(wl)
    m, I, _ = o.masses()

    def ke():
        s = o.state()
        return float(0.5 * np.sum(m[:, None] * s["v"] ** 2) + 0.5 * np.sum(I[:, None] * s["w"] ** 2))

    assert ke() == ref["kinetic_energy"][0]
    o.step(ref["every_steps"])
    assert ke() == ref["kinetic_energy"][1]
    assert len(ref["kinetic_energy"]) == 21 and ref["dt"] == wl["dt"]
