"""Discrete adjoint of the FP64 SecondOrderLeapFrog step, through the C ABI
(grains_gpu_adjoint_forward / grains_gpu_adjoint_backward) and the
torch.autograd.Function over it.

* the checkpointed forward run is the plain run, bit for bit, and the system
  is left at that end state after the reverse sweep;
* gradients do not depend on the checkpoint interval (recomputation is
  bit-identical), bit for bit;
* directional derivatives with respect to x0, v0, w0 and kn, en, muc match
  finite differences of GPU rollouts to 1e-5 relative.  The step map is only
  piecewise smooth (the damping force jumps where a contact starts or ends),
  so the differences are taken on a stencil whose rollouts all see the same
  number of contacts at every step: central, else one-sided, else with a
  smaller step.
* the 20,000-sphere inclined-plane rollout (500 steps, checkpoints every 50):
  gradients of the final centre-of-mass height and kinetic energy with
  respect to mu, en, kn and the initial velocity field against finite
  differences starting from h = 1e-6 relative, with the same event handling
  down to 1e-10.  At this size contact onsets lie about 1e-7 apart
  (relative, in en and kn) and Coulomb switches closer still; a derivative
  whose every stencil crosses a switch is compared, at 5e-3, with the 1e-6
  central difference, which averages it over the switches; event-free
  stencils are compared at 2e-3, the scatter of the differences themselves
  at this size.
"""
import numpy as np
import pytest

from pacific_amd import workloads
from tests.helpers import make_gpu

pytestmark = pytest.mark.gpu


def rollout(g, s0, n, model=None):
    if model is not None:
        g.contact_model(0, 0, *model)
    g.set_state(x=s0["x"], v=s0["v"], w=s0["w"], q=s0["q"], recompute=True)
    g.step(n)
    return g.state()


def lin_objective(rng, n):
    """J = A.x + B.v + C.w with fixed random weights"""
    return rng.normal(size=(n, 3)), rng.normal(size=(n, 3)) * 1e-3, rng.normal(size=(n, 3)) * 1e-6


def J_of(st, A, B, C):
    return float(np.sum(A * st["x"]) + np.sum(B * st["v"]) + np.sum(C * st["w"]))


def setup(wl):
    # the adjoint replays the per-step neighbour search (no Verlet skin)
    g = make_gpu(wl, record=False, skin=0.0)
    s0 = g.state()
    model = list(wl["models"][0][2:5])  # kn, en, muc
    return g, s0, model


def adjoint(g, s0, n, K, cot):
    g.set_state(x=s0["x"], v=s0["v"], w=s0["w"], q=s0["q"], recompute=True)
    g.adjoint_forward(n, K)
    end = g.state()
    grad = g.adjoint_backward(*cot)
    return end, grad


CASES = {
    "walls": dict(side=4, seed=3),
    "periodic": dict(side=4, seed=5, periodic=True, gravity=(0.0, 0.0, 0.0)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_checkpointed_forward_is_plain_forward(gpu, case):
    g, s0, _ = setup(workloads.dense_pack(**CASES[case]))
    n = 60
    plain = rollout(g, s0, n)
    rng = np.random.default_rng(1)
    end, _ = adjoint(g, s0, n, 16, lin_objective(rng, g.n))
    after = g.state()
    for k in ("x", "v", "w", "q", "F", "T"):
        assert np.array_equal(end[k], plain[k]), k
        assert np.array_equal(after[k], plain[k]), k


def test_gradient_independent_of_checkpoint_interval(gpu):
    g, s0, _ = setup(workloads.dense_pack(**CASES["walls"]))
    rng = np.random.default_rng(2)
    cot = lin_objective(rng, g.n)
    grads = [adjoint(g, s0, 45, K, cot)[1] for K in (1, 7, 45, 100)]
    for gr in grads[1:]:
        for k in ("x0", "v0", "w0", "params"):
            assert np.array_equal(gr[k], grads[0][k]), k


def rollout_events(g, s0, n, model=None):
    """Rollout one step at a time; returns the end state and, for every
    step's force evaluation, the number of contacts and of contacts on the
    Coulomb cap (the event signature: a contact that starts, ends, starts or
    stops sliding one step earlier or later changes it)."""
    if model is not None:
        g.contact_model(0, 0, *model)
    g.record_pairs(True)
    g.set_state(x=s0["x"], v=s0["v"], w=s0["w"], q=s0["q"], recompute=True)
    sig = []
    for _ in range(n):
        g.step(1)
        sig.append(g.contact_counts())
    g.record_pairs(False)
    return g.state(), tuple(sig)


def fd_derivative(Jfun, h, tries=4):
    """d/dt J(t) at 0 from J(-2h), ..., J(2h), Jfun returning (J, event
    signature).  The step map jumps where a contact starts or ends one step
    earlier or later (the damping force is not zero at zero overlap) and its
    derivative jumps where the Coulomb cap switches, so a stencil point whose
    signature differs from t = 0 lies across such an event: then the
    three-point one-sided formula is used on a side free of events, else h is
    divided by 10.  Returns (derivative, noise) or None."""
    j0, e0 = Jfun(0.0)
    for _ in range(tries):
        vals = {t: Jfun(t * h) for t in (-2, -1, 1, 2)}
        J = {t: v[0] for t, v in vals.items()}
        J[0] = j0
        clean = {t: v[1] == e0 for t, v in vals.items()}
        noise = 1e3 * 1e-16 * max(abs(v) for v in J.values()) / h
        if all(clean.values()):
            return (-J[2] + 8 * J[1] - 8 * J[-1] + J[-2]) / (12 * h), noise
        if clean[1] and clean[2]:
            return (-3 * J[0] + 4 * J[1] - J[2]) / (2 * h), noise
        if clean[-1] and clean[-2]:
            return (3 * J[0] - 4 * J[-1] + J[-2]) / (2 * h), noise
        h *= 0.1
    return None


def check_directions(label, pairs, min_used, rtol=1e-5):
    """pairs: (adjoint value, fd_derivative result)"""
    used = 0
    for a, fd in pairs:
        if fd is None:
            continue
        c, noise = fd
        used += 1
        assert abs(a - c) <= rtol * abs(c) + 4 * noise, (label, a, c)
    assert used >= min_used, (label, used)


@pytest.mark.parametrize("case", list(CASES))
def test_gradients_match_finite_differences(gpu, case):
    g, s0, model = setup(workloads.dense_pack(**CASES[case]))
    n, K = 80, 20
    rng = np.random.default_rng(3)
    A, B, C = lin_objective(rng, g.n)
    _, grad = adjoint(g, s0, n, K, (A, B, C))

    def J_state(dx=None, dv=None, dw=None, mdl=None):
        s = dict(s0)
        if dx is not None:
            s["x"] = s0["x"] + dx
        if dv is not None:
            s["v"] = s0["v"] + dv
        if dw is not None:
            s["w"] = s0["w"] + dw
        st, sig = rollout_events(g, s, n, mdl)
        return J_of(st, A, B, C), sig

    for key, h in (("x0", 1e-9), ("v0", 1e-7), ("w0", 1e-4)):
        arg = {"x0": "dx", "v0": "dv", "w0": "dw"}[key]
        pairs = []
        for d in range(4):
            D = rng.normal(size=(g.n, 3))
            pairs.append((float(np.sum(grad[key] * D)), fd_derivative(lambda t: J_state(**{arg: t * D}), h)))
        check_directions(key, pairs, 3)
    base = list(model)
    for i, name in enumerate(("kn", "en", "muc")):

        def Jp(t):
            m = list(base)
            m[i] += t
            return J_state(mdl=m)

        check_directions(name, [(grad["params"][i], fd_derivative(Jp, abs(base[i]) * 1e-6))], 1)
    g.contact_model(0, 0, *base)


def test_torch_function_matches_c_abi(gpu):
    import torch

    from pacific_amd.adjoint import rollout as torch_rollout
    g, s0, model = setup(workloads.dense_pack(**CASES["walls"]))
    n, K = 30, 10
    rng = np.random.default_rng(4)
    A, B, C = lin_objective(rng, g.n)
    _, grad = adjoint(g, s0, n, K, (A, B, C))
    x0 = torch.tensor(s0["x"], requires_grad=True)
    v0 = torch.tensor(s0["v"], requires_grad=True)
    w0 = torch.tensor(s0["w"], requires_grad=True)
    th = torch.tensor(model, dtype=torch.float64, requires_grad=True)
    g.set_state(q=s0["q"], recompute=False)
    xN, vN, wN = torch_rollout(g, x0, v0, w0, th, n, K)
    J = (torch.from_numpy(A) * xN).sum() + (torch.from_numpy(B) * vN).sum() + (torch.from_numpy(C) * wN).sum()
    J.backward()
    assert np.array_equal(x0.grad.numpy(), grad["x0"])
    assert np.array_equal(v0.grad.numpy(), grad["v0"])
    assert np.array_equal(w0.grad.numpy(), grad["w0"])
    assert np.array_equal(th.grad.numpy(), grad["params"])


# At 20,000 spheres and 500 steps the differences of the discrete map hold
# rounding noise and Coulomb-switch kinks at the 1e-4..1e-3 level even on
# event-free stencils (they scatter by that much from one step size to the
# next), so the comparison is made at 2e-3; the small systems above pin the
# adjoint itself at 1e-5.
C5_RTOL = 2e-3


def test_inclined_plane_gradients(gpu):
    """20,000 spheres onto a 25-degree plane, 500 steps, checkpoints every 50:
    d(z_com)/d and d(KE)/d (mu, en, kn, v0 along a random field)."""
    wl = workloads.inclined_plane()
    g, s0, model = setup(wl)
    n, K = 500, 50
    m, I, _ = g.masses()
    M = m.sum()

    def objectives(st):
        zc = float(np.sum(m * st["x"][:, 2]) / M)
        ke = float(0.5 * np.sum(m * np.sum(st["v"] ** 2, axis=1)) + 0.5 * np.sum(I * np.sum(st["w"] ** 2, axis=1)))
        return zc, ke

    g.set_state(x=s0["x"], v=s0["v"], w=s0["w"], q=s0["q"], recompute=True)
    g.adjoint_forward(n, K)
    end = g.state()
    zN = np.zeros((g.n, 3))
    zN[:, 2] = m / M
    g_z = g.adjoint_backward(dJ_dx=zN)
    g_ke = g.adjoint_backward(dJ_dv=m[:, None] * end["v"], dJ_dw=I[:, None] * end["w"])
    rng = np.random.default_rng(7)
    D = rng.normal(size=(g.n, 3)) * 0.01
    strict = 0
    for j, gr in enumerate((g_z, g_ke)):
        def Jv(t):
            s = dict(s0)
            s["v"] = s0["v"] + t * D
            st, sig = rollout_events(g, s, n)
            return objectives(st)[j], sig

        checks = [(f"v0[{j}]", float(np.sum(gr["v0"] * D)), Jv, 1e-6)]
        for i, name in enumerate(("kn", "en", "muc")):
            def Jp(t, i=i):
                mm = list(model)
                mm[i] += t
                st, sig = rollout_events(g, s0, n, mm)
                return objectives(st)[j], sig

            checks.append((f"{name}[{j}]", gr["params"][i], Jp, abs(model[i]) * 1e-6))
        for label, a, Jf, h in checks:
            fd = fd_derivative(Jf, h, tries=5)
            if fd is not None:
                strict += 1
                check_directions(label, [(a, fd)], 1, rtol=C5_RTOL)
            else:
                # every stencil down to 1e-10 relative crossed an event:
                # compare with the 1e-6 central difference, which averages
                # the derivative over the events it spans
                jp, jm = Jf(h)[0], Jf(-h)[0]
                c = (jp - jm) / (2 * h)
                assert abs(a - c) <= 5e-3 * abs(c), (label, a, c)
        g.contact_model(0, 0, *model)
    assert strict >= 4, strict
