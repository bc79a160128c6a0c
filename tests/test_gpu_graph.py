"""Steps replayed from captured HIP graphs against the same steps launched
one call at a time.

A batch of n LeapFrog steps runs its fused steps in chunks replayed from a
graph captured on first use (one per buffer-parity state); a batch of one
step never does.  Same kernels, same arguments, same order: every state
array is bit-identical, across the contact-search modes, both precisions,
walls, periodicity, polyhedra and the contact model with memory, and over
repeated batches that reuse the cached graphs from other parity states.
"""
import numpy as np
import pytest

from pacific_amd import MIXED, workloads
from tests.helpers import make_gpu
from tests.test_gpu_memory import with_memory

pytestmark = pytest.mark.gpu

CASES = {
    "gas_verlet": (lambda: workloads.periodic_gas(27000, sigma=0.6), {}),
    "settling_verlet": (lambda: workloads.settling_box(1000), {}),
    "dense_walls_per_step_list": (lambda: workloads.dense_pack(side=8), {"skin": 0.0}),
    "dense_periodic_cell_search": (lambda: workloads.dense_pack(side=8, periodic=True, gravity=(0.0, 0.0, 0.0)),
                                   {"nbr": 0}),
    "inclined_memory": (lambda: with_memory(workloads.inclined_plane(3000)), {}),
    "polyhedra": (lambda: workloads.polyhedra_box((5, 5, 4), seed=9, spheres=True), {}),
}


@pytest.mark.parametrize("precision", [0, MIXED])
@pytest.mark.parametrize("case", sorted(CASES))
def test_graph_replay_equals_single_steps(gpu, case, precision):
    make, kw = CASES[case]
    wl = make()
    a = make_gpu(wl, precision=precision, record=False, **kw)
    b = make_gpu(wl, precision=precision, record=False, **kw)
    # 150 = 1 + 2 x 64 + 21 fused steps; then 151 (other parities at the
    # graph's start), then 150 again (the first graph reused)
    for n in (150, 151, 150):
        a.step(n)
        for _ in range(n):
            b.step(1)
        assert a.time == b.time
        sa, sb = a.state(), b.state()
        for k in ("x", "v", "w", "wb", "q", "F", "T"):
            assert np.array_equal(sa[k], sb[k]), (case, n, k)
    ngraphs, replays, why = a.graph_stats()
    assert ngraphs >= 1 and replays == 6, (ngraphs, replays, why)
    assert b.graph_stats()[:2] == (0, 0)
