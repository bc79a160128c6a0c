"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars (from the path's contract):
  * contact records bit-exact: the set of directed (receiver, partner) pairs
    of a force evaluation equals the oracle's on the same state;
  * FP64 forces and torques within 1e-12 relative, per particle, of the
    oracle's, relative to |weight| + |contact force| of that particle (the
    summation order of a particle's contacts differs, so bits may differ);
  * integration bit-exact (x, v, w, wb) when forces agree bit for bit
    (contact-free motion), quaternions within 4 ulp (device sin/cos);
  * short rollouts within a stated tolerance; mixed FP32-force path within
    1e-5 relative force error.
"""
import numpy as np
import pytest

from pacific_amd import workloads
from tests.helpers import contact_scale, make_gpu, make_oracle, pair_set, rel_force_error

pytestmark = pytest.mark.gpu


def settled_state(wl, steps):
    """Oracle state after `steps` steps, with its forces re-evaluated at the
    end-of-step velocities (inside a step the forces see the half-step
    velocity), so that both sides evaluate the same (x, v, w)."""
    o = make_oracle(wl)
    o.step(steps)
    so = o.state()
    o.set_state(x=so["x"], v=so["v"], w=so["w"], q=so["q"], recompute=True)
    return o


def compare_forces(g, o, tol):
    sg, so = g.state(), o.state()
    _, _, wgt = o.masses()
    scale = contact_scale(so, wgt)
    eF = rel_force_error(sg["F"], so["F"], scale)
    # torque scale: lever ~ radius times the force scale
    eT = float(np.max(np.linalg.norm(sg["T"] - so["T"], axis=1) / (scale * 1.2e-3)))
    assert eF <= tol, eF
    assert eT <= tol, eT
    return eF, eT


def test_forces_and_pairs_on_settling_box(gpu):
    # 1,000 spheres in the walled box after 15,000 oracle steps (bed + falling)
    wl = workloads.settling_box(1000)
    o = settled_state(wl, 15000)
    so = o.state()
    g = make_gpu(wl)
    g.set_state(x=so["x"], v=so["v"], w=so["w"], q=so["q"], recompute=True)
    pg, po = pair_set(g.pairs()), pair_set(o.pairs())
    assert len(po) > 500
    assert pg == po
    compare_forces(g, o, 1e-12)


def test_forces_and_pairs_periodic_gas(gpu):
    wl = workloads.periodic_gas(8000, seed=4, sigma=0.5)
    o = settled_state(wl, 300)
    so = o.state()
    g = make_gpu(wl)
    g.set_state(x=so["x"], v=so["v"], w=so["w"], q=so["q"], recompute=True)
    pg, po = pair_set(g.pairs()), pair_set(o.pairs())
    assert len(po) > 20
    assert pg == po
    compare_forces(g, o, 1e-12)


def test_free_fall_steps_bit_exact(gpu):
    # before the first contact every force is the weight: bit-exact state
    wl = workloads.settling_box(1000)
    o, g = make_oracle(wl), make_gpu(wl)
    for n in (1, 5, 20):
        o.step(n)
        g.step(n)
        so, sg = o.state(), g.state()
        for k in ("x", "v", "w", "wb", "F"):
            assert np.array_equal(sg[k], so[k]), k
        assert np.array_equal(sg["q"], so["q"])


def test_spinning_particles_quaternion(gpu):
    # free rotation with w != 0: x, v, w, wb bit-exact; q within a few ulp
    rng = np.random.default_rng(7)
    wl = workloads.settling_box(200)
    wl["w"] = rng.normal(0, 50.0, (200, 3))
    wl["gravity"] = (0.0, 0.0, 0.0)
    o, g = make_oracle(wl), make_gpu(wl)
    o.step(300)
    g.step(300)
    so, sg = o.state(), g.state()
    for k in ("x", "v"):
        assert np.array_equal(sg[k], so[k]), k
    assert np.max(np.abs(sg["q"] - so["q"])) <= 4 * np.finfo(float).eps
    assert np.max(np.abs(sg["w"] - so["w"])) <= 1e-13 * np.max(np.abs(so["w"]))


def test_short_rollout_settling(gpu):
    # 3,000 steps through the first impacts: same contact records at the end,
    # state within 1e-9 relative (contact sums differ in rounding order)
    wl = workloads.settling_box(1000)
    o, g = make_oracle(wl), make_gpu(wl)
    o.step(3000)
    g.step(3000)
    so, sg = o.state(), g.state()
    assert np.max(np.abs(sg["x"] - so["x"])) < 1e-9 * 0.06
    vs = np.max(np.abs(so["v"]))
    assert np.max(np.abs(sg["v"] - so["v"])) < 1e-7 * vs
    assert pair_set(g.pairs()) == pair_set(o.pairs())


@pytest.mark.parametrize("scheme", [1, 2, 3])
def test_other_integrators(gpu, scheme):
    wl = workloads.settling_box(300)
    o, g = make_oracle(wl, scheme=scheme), make_gpu(wl, scheme=scheme)
    o.step(50)
    g.step(50)
    so, sg = o.state(), g.state()
    assert np.array_equal(sg["x"], so["x"])
    assert np.array_equal(sg["v"], so["v"])


def test_mixed_precision_force_error(gpu):
    wl = workloads.settling_box(1000)
    o = settled_state(wl, 15000)
    so = o.state()
    g = make_gpu(wl, precision=1)
    g.set_state(x=so["x"], v=so["v"], w=so["w"], q=so["q"], recompute=True)
    assert pair_set(g.pairs()) == pair_set(o.pairs())
    compare_forces(g, o, 1e-5)


def test_device_errors(gpu):
    from pacific_amd import GrainsError
    wl = workloads.settling_box(10)
    wl["v"][3] = (0.0, 0.0, -10.0)
    g = make_gpu(wl)
    with pytest.raises(GrainsError) as e:
        g.step(1)
    assert e.value.kind == "MotionError" and e.value.ids[0] == 4
