"""Shared helpers of the parity tests: build the same workload on the HIP
path and on the CPU oracle, and compare the results."""
import numpy as np

from oracle.oracle import Oracle
from pacific_amd import workloads


def make_oracle(wl, scheme=0, fluid_density=0.0, collision=None):
    """collision: keyword arguments of Oracle.collision (GJK_SV, tolerance,
    acceleration, bounding volume), set before the initial forces."""
    o = Oracle(wl["origin"], wl["max"], wl["periodic"], wl["cellsize_factor"], wl["gravity"], fluid_density,
               wl["dt"], scheme)
    workloads.build(o, wl)
    if collision:
        o.collision(**collision)
    o.init()
    return o


def make_gpu(wl, scheme=0, precision=0, fluid_density=0.0, record=True, nbr=None, skin=None, stats=False,
             collision=None):
    """nbr: candidates per particle (0: cell search in the force sweep);
    skin: Verlet skin (0: per-step list); None: the library defaults;
    stats: collect contact statistics from the first force evaluation;
    collision: keyword arguments of GrainsGPU.collision."""
    from pacific_amd import GrainsGPU
    g = GrainsGPU(wl["origin"], wl["max"], wl["periodic"], wl["cellsize_factor"], wl["gravity"], fluid_density,
                  wl["dt"], scheme, precision)
    workloads.build(g, wl)
    if collision:
        g.collision(**collision)
    if nbr is not None or skin is not None:
        g.neighbor_list(32 if nbr is None else nbr, -1.0 if skin is None else skin)
    if record:
        g.record_pairs(True)
    if stats:
        g.contact_stats(True)
    g.init()
    return g


def pair_set(p):
    return set(map(tuple, np.asarray(p).tolist()))


def rel_force_error(F, Fref, scale):
    """max over particles of |F - Fref| / scale (scale = per-particle sum of
    contact-force magnitudes plus the weight)"""
    err = np.linalg.norm(F - Fref, axis=1)
    return float(np.max(err / scale))


def contact_scale(state, weight, pairs=None):
    """Per-particle force scale: |weight| + |F_contact| with a floor."""
    Fc = state["F"] - weight
    s = np.linalg.norm(weight, axis=1) + np.linalg.norm(Fc, axis=1)
    return np.maximum(s, 1e-30)


def min_image_pairs(x, r, L=None):
    """Brute-force directed contact pairs (i+1, j+1) with |xj - xi| - (ri + rj) < 0,
    minimum image on periodic axes (L[a] > 0)."""
    n = len(x)
    out = []
    for i in range(n):
        d = x - x[i]
        if L is not None:
            for a in range(3):
                if L[a] > 0:
                    d[:, a] -= L[a] * np.round(d[:, a] / L[a])
        dist = np.sqrt(np.sum(d * d, axis=1)) - (r + r[i])
        js = np.nonzero(dist < 0)[0]
        out += [(i + 1, j + 1) for j in js if j != i]
    return set(out)
