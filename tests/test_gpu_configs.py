"""Parity of the HIP path with the oracle on the configurations and cases that
the other GPU tests leave out: the polyhedra silo discharging from rest (FP64
and mixed), the mixed periodic gas, a batched member once its bed has formed,
the rolling-resistance moment (kr != 0) and the device's ContactError path.

Bars: contact pair sets equal; FP64 forces and torques within 1e-12 per
particle relative to |weight| + |contact force|; mixed forces within 1e-5 per
particle relative to the sum of the magnitudes of that particle's contact
forces (sum |f_contact| of the oracle's evaluation); short rollouts within the
stated tolerances; unrecoverable overlaps stop the run with ContactError and
the ids of the two bodies.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, OracleError
from pacific_amd import GrainsBatch, GrainsError, GrainsGPU, workloads
from tests.helpers import contact_scale, make_gpu, make_oracle, pair_set, rel_force_error

pytestmark = pytest.mark.gpu

RHO = 2500.0


def mixed_error(sg, o):
    """Mixed-precision force error against the oracle's FP64 evaluation, per
    particle with contacts: |F - F_oracle| relative to (a) the scale of the
    terms its contact forces are made of, sum over contacts of kn |pen| +
    (|gamman| + |gammat|) |v_rel| (oracle term_scales) -- FP32 rounding is
    relative to these operands -- and (b) the sum of the magnitudes of its
    contact forces, sum |f_contact| (oracle force_sums), which can be far
    smaller where the spring and the dashpot cancel at the end of a contact.
    Particles without contacts must carry the oracle's force exactly.
    Returns (max of a, max of b, mean of b, 99.9th percentile of b)."""
    so = o.state()
    fs, ts = o.force_sums(), o.term_scales()
    err = np.linalg.norm(sg["F"] - so["F"], axis=1)
    touch = fs > 0
    assert np.all(err[~touch] == 0.0)
    if not np.any(touch):
        return 0.0, 0.0, 0.0, 0.0
    a, b = err[touch] / ts[touch], err[touch] / fs[touch]
    return float(a.max()), float(b.max()), float(b.mean()), float(np.quantile(b, 0.999))


def assert_mixed_bars(errs):
    term_max, fsum_max, fsum_mean, fsum_p999 = errs
    assert term_max <= 1e-5, errs
    assert fsum_mean <= 1e-6, errs
    assert fsum_p999 <= 1e-4, errs


def fp64_errors(sg, o, lever):
    so = o.state()
    _, _, wgt = o.masses()
    scale = contact_scale(so, wgt)
    eF = rel_force_error(sg["F"], so["F"], scale)
    eT = float(np.max(np.linalg.norm(sg["T"] - so["T"], axis=1) / (scale * lever)))
    return eF, eT


def to_oracle(wl, sg, basis=None, x=None):
    """Oracle of workload wl placed in the state sg (forces re-evaluated)."""
    o = make_oracle(wl)
    if basis is not None:
        o.set_basis(basis)
    o.set_state(x=sg["x"] if x is None else x, v=sg["v"], w=sg["w"], q=sg["q"], recompute=True)
    return o


# ---------------------------------------------------------------------------
# C3: polyhedra silo discharging from rest

def small_silo():
    # 2,420 bodies (thirds cube / tetrahedron / icosahedron, 20 layers of
    # 11 x 11) in a 39.6 mm silo with the 12 mm orifice and its collecting
    # floor 30 mm below, from rest
    return workloads.silo(n=2420, width=0.04)


@pytest.fixture(scope="module")
def silo_state(gpu):
    wl = small_silo()
    g = make_gpu(wl)
    # the lattice falls onto the floor plates, the bed builds up while the
    # bodies above the orifice discharge into the collecting bin
    g.step(24000)
    sg = g.state()
    return wl, sg, g.basis().reshape(-1, 9)


def test_silo_fp64_pairs_forces_and_rollout(gpu, silo_state):
    wl, s0, B0 = silo_state
    o = to_oracle(wl, s0, B0)
    g = make_gpu(wl)
    g.set_basis(B0)
    g.set_state(x=s0["x"], v=s0["v"], w=s0["w"], q=s0["q"], recompute=True)
    po, pg = pair_set(o.pairs()), pair_set(g.pairs())
    assert len(po) > 600
    assert pg == po
    eF, eT = fp64_errors(g.state(), o, 1.8e-3)
    assert eF <= 1e-12, eF
    assert eT <= 1e-12, eT
    # 300 steps of the discharge from the same state.  The contact sums run
    # in another order (1e-16 apart), and a face lying flat on another body
    # has tied support vertices: an ulp in the search direction moves the
    # witness point across the face, so single bodies drift apart while the
    # flow as a whole agrees.  Bars: the median body within 1e-12 m, every
    # body within 1e-5 m (1/200 of a body), contact records 95% alike (97%
    # measured: records of bodies whose contacts open or close in the steps
    # where they drifted apart).
    o.step(300)
    g.step(300)
    so, sg = o.state(), g.state()
    dx = np.linalg.norm(sg["x"] - so["x"], axis=1)
    assert np.median(dx) < 1e-12, np.median(dx)
    assert dx.max() < 1e-5, dx.max()
    pg, po = pair_set(g.pairs()), pair_set(o.pairs())
    assert len(pg & po) >= 0.95 * len(pg | po), (len(pg & po), len(pg | po))


def test_silo_mixed_forces(gpu, silo_state):
    wl, s0, B0 = silo_state
    o = to_oracle(wl, s0, B0)
    g = make_gpu(wl, precision=1)
    g.set_basis(B0)
    g.set_state(x=s0["x"], v=s0["v"], w=s0["w"], q=s0["q"], recompute=True)
    assert pair_set(g.pairs()) == pair_set(o.pairs())
    assert_mixed_bars(mixed_error(g.state(), o))


# ---------------------------------------------------------------------------
# C2 at a size the oracle covers: mixed periodic gas

def test_mixed_periodic_gas_against_oracle(gpu):
    wl = workloads.periodic_gas(8000, seed=4, sigma=0.5)
    o = make_oracle(wl)
    o.step(300)
    so = o.state()
    o.set_state(x=so["x"], v=so["v"], w=so["w"], q=so["q"], recompute=True)
    g = make_gpu(wl, precision=1)
    g.set_state(x=so["x"], v=so["v"], w=so["w"], q=so["q"], recompute=True)
    po = pair_set(o.pairs())
    assert len(po) > 20
    assert pair_set(g.pairs()) == po
    assert_mixed_bars(mixed_error(g.state(), o))


# ---------------------------------------------------------------------------
# C4: a batched member once its bed has formed

def test_batch_member_against_oracle_after_bed_forms(gpu):
    ens = workloads.ensemble(nsys=3, seed=11, nmin=150, nmax=900)

    def unconfigured(w):
        g = GrainsGPU(w["origin"], w["max"], w["periodic"], w["cellsize_factor"], w["gravity"], 0.0, w["dt"], 0, 0)
        workloads.build(g, w)
        return g

    batch = GrainsBatch([unconfigured(w) for w in ens])
    # the beds form from step ~2,000 on; at 8,000 each member holds hundreds
    # of contact records
    batch.step(8000, chunk=500)
    for i, wl in enumerate(ens):
        sb = batch.state(i)
        m, I, _ = make_oracle(wl).masses()
        dt = wl["dt"]
        # the member's torsor was evaluated at (x, v - dt/2 F/m, w - dt/2 T/I)
        # (LeapFrog: forces at the half-step velocity); rebuild that velocity
        vh = sb["v"] - 0.5 * dt * sb["F"] / m[:, None]
        wh = sb["w"] - 0.5 * dt * sb["T"] / I[:, None]
        o = make_oracle(wl)
        o.set_state(x=sb["x"], v=vh, w=wh, q=sb["q"], recompute=True)
        assert len(o.pairs()) > 300, i
        eF, eT = fp64_errors(sb, o, 1.2e-3)
        assert eF <= 1e-12, (i, eF)
        assert eT <= 1e-12, (i, eT)


# ---------------------------------------------------------------------------
# Rolling resistance (kr != 0), HookeContactForceModel.cpp:95-111

@pytest.mark.parametrize("precision", [0, 1])
def test_rolling_resistance_parity_dense_pack(gpu, precision):
    wl = workloads.dense_pack(6, seed=5, sigma=0.05)
    wl["models"] = [(0, 0, 1e4, 0.8, 0.4, 0.02, -1.0)]
    wl["w"] = np.random.default_rng(9).normal(0, 20.0, (len(wl["x"]), 3))
    o = make_oracle(wl)
    g = make_gpu(wl, precision=precision)
    so = o.state()
    g.set_state(x=so["x"], v=so["v"], w=so["w"], q=so["q"], recompute=True)
    po = pair_set(o.pairs())
    assert len(po) > 500
    assert pair_set(g.pairs()) == po
    sg = g.state()
    if precision == 0:
        eF, eT = fp64_errors(sg, o, 1.2e-3)
        assert eF <= 1e-12 and eT <= 1e-12, (eF, eT)
    else:
        assert_mixed_bars(mixed_error(sg, o))
    # the rolling moment carries a visible share of the torque
    o2 = make_oracle(dict(wl, models=[(0, 0, 1e4, 0.8, 0.4, 0.0, -1.0)]))
    o2.set_state(x=so["x"], v=so["v"], w=so["w"], q=so["q"], recompute=True)
    dT = np.abs(so["T"] - o2.state()["T"]).max()
    assert dT > 1e-3 * np.abs(so["T"]).max()


@pytest.mark.parametrize("precision", [0, 1])
def test_rolling_resistance_closed_form(gpu, precision):
    # frictionless head-on collision of two spinning spheres: the spin
    # difference drops by 2 kr Req r J (w0 + w1) / I and w0 + w1 is conserved
    # (the derivation is in tests/test_oracle_closed_forms.py)
    kn, en, r, kr = 1e4, 0.9, 1e-3, 0.6
    m = RHO * 4 / 3 * math.pi * r ** 3
    inertia = 0.4 * m * r ** 2
    mstar = m / 2
    w0n = math.sqrt(kn / mstar)
    beta = math.log(en) / math.sqrt(math.pi ** 2 + math.log(en) ** 2)
    tc = math.pi / math.sqrt(w0n ** 2 - (beta * w0n) ** 2)
    dt = tc / 2000
    u, w0, w1 = 0.2, 100.0, 40.0
    g = GrainsGPU((0, 0, 0), (0.02, 0.02, 0.02), gravity=(0, 0, 0), dt=dt, precision=precision)
    g.contact_model(0, 0, kn, en, 0.0, kr=kr)
    g.add_spheres(r, RHO, 5e-5, 0, [[0.009 - 1e-6, 0.01, 0.01], [0.011, 0.01, 0.01]],
                  [[u / 2, 0, 0], [-u / 2, 0, 0]], [[0, 0, w0], [0, 0, w1]])
    g.init()
    g.step(4000)
    s = g.state()
    J = m * (u / 2 - s["v"][0, 0])
    drop = (w0 - w1) - (s["w"][0, 2] - s["w"][1, 2])
    expected = 2 * kr * (r / 2) * r * J * (w0 + w1) / inertia
    assert drop == pytest.approx(expected, rel=3e-3)
    assert s["w"][0, 2] + s["w"][1, 2] == pytest.approx(w0 + w1, rel=1e-6 if precision else 1e-12)
    if precision == 0:
        o = Oracle((0, 0, 0), (0.02, 0.02, 0.02), gravity=(0, 0, 0), dt=dt)
        o.contact_model(0, 0, kn, en, 0.0, kr=kr)
        o.add_spheres(r, RHO, 5e-5, 0, [[0.009 - 1e-6, 0.01, 0.01], [0.011, 0.01, 0.01]],
                      [[u / 2, 0, 0], [-u / 2, 0, 0]], [[0, 0, w0], [0, 0, w1]])
        o.init()
        o.step(4000)
        so = o.state()
        assert np.abs(s["w"] - so["w"]).max() <= 1e-12 * w0
        assert np.abs(s["x"] - so["x"]).max() <= 1e-15


# ---------------------------------------------------------------------------
# ContactError on the device (Particle.cpp:477-545, RigidBodyWithCrust.cpp:534-590)

def two_spheres(gap_centres, backend):
    # radius 1 mm, crust 5e-5: overlaps beyond the crust sum (centres closer
    # than 1.9 mm) are recovered by shrinking both spheres by ten crusts
    # (radius 0.5 mm): recoverable while the centres are more than 1 mm apart
    sysm = backend((0, 0, 0), (0.02, 0.02, 0.02), gravity=(0, 0, 0), dt=1e-5)
    sysm.contact_model(0, 0, 1e4, 0.8, 0.4)
    sysm.add_spheres(1e-3, RHO, 5e-5, 0, [[0.01, 0.01, 0.01], [0.01 + gap_centres, 0.01, 0.01]],
                     [[0.01, 0, 0], [0, 0.02, 0]], [[0, 0, 5.0], [3.0, 0, 0]])
    return sysm


def sphere_on_floor(z, backend):
    # floor 0.5 mm thick, top face at z = 0; the recovery shrinks the floor's
    # half thickness to 0.125 mm and the sphere to 0.5 mm: recoverable while
    # the centre is more than 0.375 mm above the top face
    sysm = backend((0, 0, -0.002), (0.02, 0.02, 0.02), gravity=(0, 0, 0), dt=1e-5)
    sysm.contact_model(0, 0, 1e4, 0.8, 0.4)
    sysm.add_spheres(1e-3, RHO, 5e-5, 0, [[0.01, 0.01, z]], [[0.02, 0, 0]], [[0, 4.0, 0]])
    sysm.add_box((0.021, 0.021, 0.0005), (0.01, 0.01, -0.00025))
    return sysm


def two_cubes(gap_centres, backend):
    # cubes of 2 mm edge face to face along x: the crust-shrunk cubes (half
    # edge 0.95 mm) meet for centres closer than 1.9 mm; the recovery shrinks
    # them to half edge 0.5 mm: recoverable while the centres are more than
    # 1 mm apart
    sysm = backend((0, 0, 0), (0.02, 0.02, 0.02), gravity=(0, 0, 0), dt=5e-6)
    sysm.contact_model(0, 0, 1e4, 0.8, 0.4)
    c = sysm.add_shape(1, L=(2e-3, 2e-3, 2e-3), crust=5e-5)
    sysm.add_particles(c, RHO, 0, [[0.01, 0.01, 0.01], [0.01 + gap_centres, 0.0101, 0.0099]],
                       [[0.01, 0, 0], [0, 0.02, 0]], [[0, 0, 5.0], [3.0, 0, 0]],
                       np.tile(np.eye(3).reshape(1, 9), (2, 1)))
    return sysm


def _gpu(*a, **k):
    return GrainsGPU(*a, **k)


def _orc(*a, **k):
    return Oracle(*a, **k)


@pytest.mark.parametrize("case,arg", [("spheres", 1.5e-3), ("floor", 0.8e-3), ("cubes", 1.5e-3)])
def test_contact_error_recovered_like_the_oracle(gpu, case, arg):
    build = {"spheres": two_spheres, "floor": sphere_on_floor, "cubes": two_cubes}[case]
    o = build(arg, _orc)
    o.init()
    assert o.stats()["recovered"] >= 1
    g = build(arg, _gpu)
    g.record_pairs(True)
    g.init()
    assert g.recoveries() >= 1
    so, sg = o.state(), g.state()
    assert pair_set(g.pairs()) == pair_set(o.pairs())
    # the recovered contact imposes an overlap of one crust sum: |F_n| = kn 1e-4
    assert np.linalg.norm(so["F"][0]) > 0.5
    assert np.abs(sg["F"] - so["F"]).max() <= 1e-12 * np.abs(so["F"]).max()
    assert np.abs(sg["T"] - so["T"]).max() <= 1e-12 * np.abs(so["F"]).max() * 2e-3


@pytest.mark.parametrize("case,arg,ids", [("spheres", 0.8e-3, {1, 2}), ("floor", 0.2e-3, {1, -1}),
                                          ("cubes", 0.8e-3, {1, 2})])
def test_contact_error_unrecoverable(gpu, case, arg, ids):
    build = {"spheres": two_spheres, "floor": sphere_on_floor, "cubes": two_cubes}[case]
    o = build(arg, _orc)
    with pytest.raises(OracleError) as eo:
        o.init()
    assert eo.value.kind == "ContactError"
    g = build(arg, _gpu)
    with pytest.raises(GrainsError) as e:
        g.init()
        g.step(1)
    assert e.value.kind == "ContactError"
    assert set(e.value.ids) == ids
