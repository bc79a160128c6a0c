"""Batched ensembles (flat mode, the default: one launch per phase over the
particles of every member; persistent mode: one launch, one workgroup per
member): every member evolves bit for bit like the same system stepped on
its own, in both modes, through direct launches and through replays of the
captured hipGraph."""
import numpy as np
import pytest

from pacific_amd import GrainsBatch, workloads
from tests.helpers import make_gpu, make_oracle

pytestmark = pytest.mark.gpu


def members(n=6, seed=3):
    ens = workloads.ensemble(nsys=n, seed=seed, nmin=150, nmax=900)
    return ens


def unconfigured(wl, precision=0):
    from pacific_amd import GrainsGPU
    g = GrainsGPU(wl["origin"], wl["max"], wl["periodic"], wl["cellsize_factor"], wl["gravity"], 0.0, wl["dt"], 0,
                  precision)
    workloads.build(g, wl)
    return g


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("shape", [0, 256])
def test_members_match_single_systems(gpu, precision, shape):
    ens = members()
    batch = GrainsBatch([unconfigured(w, precision) for w in ens], precision=precision)
    batch.set_block(shape)
    batch.step(1500, chunk=500)
    for i, wl in enumerate(ens):
        g = make_gpu(wl, precision=precision, record=False)
        g.step(1500)
        sb, sg = batch.state(i), g.state()
        for k in ("x", "v", "w", "q", "F"):
            assert np.array_equal(sb[k], sg[k]), (i, k)


@pytest.mark.parametrize("shape", [0, 256])
def test_graph_replay_equals_direct_launch(gpu, shape):
    ens = members(4, seed=8)
    b1 = GrainsBatch([unconfigured(w) for w in ens])
    b2 = GrainsBatch([unconfigured(w) for w in ens])
    b1.set_block(shape)
    b2.set_block(shape)
    b1.step(600, chunk=0)
    b2.step(600, chunk=150)
    for i in range(len(ens)):
        s1, s2 = b1.state(i), b2.state(i)
        assert np.array_equal(s1["x"], s2["x"]) and np.array_equal(s1["v"], s2["v"])


def test_member_against_oracle(gpu):
    wl = members(3, seed=11)[1]
    batch = GrainsBatch([unconfigured(w) for w in members(3, seed=11)])
    o = make_oracle(wl)
    o.step(40)
    batch.step(40)
    so, sb = o.state(), batch.state(1)
    # free fall so far: bit-exact with the oracle
    assert np.array_equal(sb["x"], so["x"]) and np.array_equal(sb["v"], so["v"])


def test_autotune_picks_a_candidate(gpu):
    batch = GrainsBatch([unconfigured(w) for w in members(8, seed=2)])
    best, ms = batch.autotune(10)
    assert best in (0, 64, 128, 192, 256, 320, 384, 448, 512) and len(ms) == 9
    assert all(t > 0 for t in ms.values())
    assert ms[best] == min(ms.values())


@pytest.mark.parametrize("precision", [0, 1])
def test_flat_mode_matches_persistent_and_single(gpu, precision):
    # the flat launches run the persistent kernel's device functions in the
    # same order: members bit for bit as there and as the system alone,
    # through direct launches and graph replays, with an odd step count (the
    # velocity buffers end swapped and are copied back)
    ens = members(5, seed=4)
    bf = GrainsBatch([unconfigured(w, precision) for w in ens], precision=precision)
    bp = GrainsBatch([unconfigured(w, precision) for w in ens], precision=precision)
    bf.set_block(0)
    bp.set_block(256)
    bf.step(1201, chunk=400)
    bp.step(1201, chunk=400)
    bf.step(3)
    bp.step(3)
    for i, wl in enumerate(ens):
        g = make_gpu(wl, precision=precision, record=False, nbr=0)
        g.step(1204)
        sf, sp, sg = bf.state(i), bp.state(i), g.state()
        for k in ("x", "v", "w", "q", "F", "T"):
            assert np.array_equal(sf[k], sp[k]), (i, k)
            assert np.array_equal(sf[k], sg[k]), (i, k)
