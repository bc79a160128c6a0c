"""grains_gpu_force_sums_noalloc (the entry point that allocates nothing:
its per-particle partial sums live in a static device array) against
grains_gpu_force_sums (a buffer the library allocates), bit for bit, on two
calls in a row, the second on a smaller system."""
import ctypes

import numpy as np
import pytest

from pacific_amd import workloads
from pacific_amd.grains import lib
from tests.helpers import make_gpu

pytestmark = pytest.mark.gpu


def noalloc_sums(g):
    import torch
    out = torch.full((g.n,), -1.0, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    code = lib().grains_gpu_force_sums_noalloc(g.h, ctypes.c_void_p(out.data_ptr()))
    assert code == 0, lib().grains_gpu_last_error(g.h)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_force_sums_match_allocating_entry(gpu):
    # a settled 1,000-sphere box, then a smaller dense pack with rolling
    # resistance: the second call overwrites only part of the shared array
    wl1 = workloads.settling_box(1000)
    g1 = make_gpu(wl1, record=False)
    g1.step(15000)
    a1 = noalloc_sums(g1)
    b1 = g1.force_sums()
    assert np.count_nonzero(b1) > 500
    assert np.array_equal(a1, b1)
    wl2 = workloads.dense_pack(5, seed=2)
    wl2["models"] = [(0, 0, 1e4, 0.8, 0.4, 0.02, -1.0)]
    g2 = make_gpu(wl2, precision=1, record=False)
    a2 = noalloc_sums(g2)
    b2 = g2.force_sums()
    assert len(a2) == 125 < len(a1)
    assert np.count_nonzero(b2) == 125
    assert np.array_equal(a2, b2)
    # and the first system again, after the smaller one
    assert np.array_equal(noalloc_sums(g1), b1)
