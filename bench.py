#!/usr/bin/env python
"""Throughput of the Grains3D DEM time step on one MI355X.

Workload: 1,000,000 polydisperse spheres, radii U[0.8, 1.2] mm, triply
periodic cube at solid fraction 0.3, Gaussian velocities (sigma 0.1 m/s), no
gravity, Hooke en = 0.8, mu = 0.3, kn = 1e4, dt = 1e-5 (homogeneous cooling
granular gas), SecondOrderLeapFrog.  The headline value is the mixed path
(FP32 force/torque arithmetic, FP64 positions, quaternions and velocities);
the plain FP64 path is measured on the same state and reported beside it.

A step is one full time step (kick, drift, re-binning, contact detection,
forces, kick) over every particle; value = N * K / (elapsed time of K steps),
state resident in HBM, timed with HIP events on the library's stream.

Prints one JSON line.  Only N = 1 GPU is supported.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle-step: state read once and written once per
# step.  FP64: x, v, w, q read+write (13 f64 x 8 B x 2 = 208 B), radius read
# (8 B), F and T carried between the half kicks (6 f64 read+write = 96 B),
# cell key write+read (8 B) -> 320 B; mixed: F, T in FP32 -> 272 B; plus 16 B
# per contact pair (one int32 pair written and read).
B_ALG = {"fp64": 320.0, "mixed": 272.0}
B_PAIR = 16.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


NBR_CAP = 32  # neighbour candidates per particle (0: the force sweep searches the 27 cells)
SKIN = -1.0  # Verlet skin (< 0: the library default, a tenth of the cell size; 0: per-step list)


def build_gpu(wl, precision, skin=None, collision=None):
    from pacific_amd import GrainsGPU, workloads
    g = GrainsGPU(wl["origin"], wl["max"], wl["periodic"], wl["cellsize_factor"], wl["gravity"], 0.0, wl["dt"],
                  0, precision)
    workloads.build(g, wl)
    g.neighbor_list(NBR_CAP, SKIN if skin is None else skin)
    if collision:
        g.collision(**collision)
    g.init()
    return g


def contacts_per_particle(g):
    g.record_pairs(True)
    g.compute_forces()
    n_rec = len(g.pairs())  # directed records = 2 per particle-particle contact
    g.record_pairs(False)
    return n_rec / 2.0 / g.n


def cpu_baseline(wl, steps):
    """The CPU oracle (serial FP64 restatement of the reference step, one
    core) on the same workload for a bounded number of steps."""
    from oracle.oracle import Oracle
    from pacific_amd import workloads
    o = Oracle(wl["origin"], wl["max"], wl["periodic"], wl["cellsize_factor"], wl["gravity"], 0.0, wl["dt"], 0)
    workloads.build(o, wl)
    o.init()
    t0 = time.perf_counter()
    o.step(steps)
    t1 = time.perf_counter()
    return wl["radius"].size * steps / (t1 - t0), t1 - t0


def host_cpu():
    """CPU model and the cores this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": len(os.sched_getaffinity(0))}


def _oracle_member(job):
    """One ensemble system on the CPU oracle (a worker process)."""
    from oracle.oracle import Oracle
    from pacific_amd import workloads
    wl, steps = job
    o = Oracle(wl["origin"], wl["max"], wl["periodic"], wl["cellsize_factor"], wl["gravity"], 0.0, wl["dt"], 0)
    workloads.build(o, wl)
    o.init()
    o.step(steps)
    return wl["radius"].size * steps


def ensemble_cpu_baseline(nsys, steps, cores):
    """The CPU oracle over the first nsys systems of the batched ensemble,
    one system per worker process on `cores` cores (the reference runs one
    system per process).  Runs before anything touches the GPU: the workers
    are forked."""
    import multiprocessing as mp
    from pacific_amd import workloads
    ens = workloads.ensemble(512, seed=5)[:nsys]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        work = sum(pool.map(_oracle_member, [(wl, steps) for wl in ens], chunksize=1))
    dt = time.perf_counter() - t0
    return work / dt, dt


def pair_contacts(x, r, box=None):
    """Particle-particle contacts (|x_i - x_j| < r_i + r_j) of one state,
    minimum image in a periodic cube of side box."""
    from scipy.spatial import cKDTree
    if box is not None:
        x = np.mod(x, box)
        tree = cKDTree(x, boxsize=box)
    else:
        tree = cKDTree(x)
    pr = tree.query_pairs(2.0 * float(r.max()), output_type="ndarray")
    if len(pr) == 0:
        return 0
    d = x[pr[:, 0]] - x[pr[:, 1]]
    if box is not None:
        d -= box * np.round(d / box)
    return int(np.count_nonzero(np.sqrt(np.sum(d * d, axis=1)) < r[pr[:, 0]] + r[pr[:, 1]]))


def batched_leg(args, prec):
    """512 independent settling boxes (sizes log-uniform in [200, 5000],
    per-system mu and en), each hipGraph replay a chunk of steps of the
    batch.  The ensemble settles untimed until every layer has landed and the
    beds form, the autotuner picks the launch shape (flat mode, or the
    persistent launch and its workgroup size) on that contact-bearing state,
    then 5,000 steps are timed; contacts per particle from the member states
    at the end."""
    from pacific_amd import GrainsBatch, GrainsGPU, workloads
    ens = workloads.ensemble(512, seed=5)
    members = []
    for wl in ens:
        g = GrainsGPU(wl["origin"], wl["max"], wl["periodic"], wl["cellsize_factor"], wl["gravity"], 0.0, wl["dt"],
                      0, prec)
        workloads.build(g, wl)
        members.append(g)
    b = GrainsBatch(members, precision=prec)
    del members
    chunk = args.batch_chunk
    t0 = time.perf_counter()
    for done in range(0, args.batch_settle, 5000):
        b.step(min(5000, args.batch_settle - done), chunk=chunk)
        log(f"bench: ensemble settle {min(done + 5000, args.batch_settle)}/{args.batch_settle} steps "
            f"({time.perf_counter() - t0:.1f} s)")
    best, tune = b.autotune(20)
    ms = b.step_timed(args.batch_steps, chunk=chunk)
    ntot = b.n_total
    value = ntot * args.batch_steps / (ms / 1e3)
    ncont = 0
    for i, wl in enumerate(ens):
        ncont += pair_contacts(b.state(i)["x"], np.asarray(wl["radius"], float) * np.ones(len(wl["x"])))
    cbar = ncont / ntot
    t_end = args.batch_settle + len(tune) * 20 + args.batch_steps
    shape = "flat" if best == 0 else f"persistent, workgroup {best}"
    log(f"bench: batch tuner (ms per {20} steps): " + ", ".join(f"{'flat' if k == 0 else k}: {v:.2f}"
                                                             for k, v in tune.items()))
    log(f"bench: batched ensemble 512 systems, {ntot} particles, {shape}: {value:.4g} particle-steps/s "
        f"(steps {t_end - args.batch_steps}-{t_end}, particle contacts/particle {cbar:.3f} at the end)")
    # one system of the same density and history for comparison: the
    # ensemble's tallest box (26 lattice layers) widened to 200,000 spheres,
    # default search (Verlet list), same untimed settle, 2,000 steps timed
    from pacific_amd import FP64
    side = int(np.ceil(np.sqrt(200_000 / 26)))
    wl1 = workloads.settling_box(200_000, seed=77, box=(side * 2.5e-3, side * 2.5e-3, 26 * 3.75e-3))
    g1 = build_gpu(wl1, FP64 if prec == 0 else prec)
    g1.step(args.batch_settle)
    g1.synchronize()
    ms1 = g1.step_timed(2000)
    single = 200_000 * 2000 / (ms1 / 1e3)
    c1 = contacts_per_particle(g1)
    del g1
    log(f"bench: one dense box of 200,000 spheres after the same settle: {single:.4g} particle-steps/s "
        f"(particle contacts/particle {c1:.3f}); batched / single {value / single:.3f}")
    return {"batched_value": value,
            "batched_single_dense_value": single,
            "batched_single_dense_config": {"particles": 200_000, "box_mm": [side * 2.5, side * 2.5, 26 * 3.75],
                                            "settle_steps": args.batch_settle, "timed_steps": 2000,
                                            "particle_contacts_per_particle": c1, "search": "Verlet list"},
            "batched_config": {"systems": 512, "particles": ntot, "precision": "fp64" if prec == 0 else "mixed",
                               "untimed_settle_steps": args.batch_settle, "autotune_trial_steps": 20,
                               "timed_steps": args.batch_steps, "timed_window": [t_end - args.batch_steps, t_end],
                               "particle_contacts_per_particle_at_end": cbar, "launch_shape": shape,
                               "graph_chunk": chunk,
                               "autotune_ms": {("flat" if k == 0 else str(k)): v for k, v in tune.items()}}}


def kinetic_energy(g, m, I):
    s = g.state()
    return float(0.5 * np.sum(m[:, None] * s["v"] ** 2) + 0.5 * np.sum(I[:, None] * s["w"] ** 2)), s


def full_run_leg(args, wl):
    """The whole cooling run (args.full_run steps from the lattice start) in
    each precision, timed in chunks of 1,000 steps with the kinetic energy
    between chunks (outside the timed chunks) and the particle contacts per
    particle every 5,000 steps; Haff's law T0 / (1 + t/tau)^2 fitted to the
    kinetic energy after the first 5,000 steps (KE^-1/2 linear in t)."""
    from pacific_amd import FP64, MIXED
    out = {"steps": args.full_run, "chunk": 1000}
    box = float(wl["max"][0] - wl["origin"][0])
    r = np.asarray(wl["radius"], float)
    curves = {}
    for prec, nm in ((FP64, "fp64"), (MIXED, "mixed")):
        g = build_gpu(wl, prec)
        m, I, _ = g.masses()
        ke0, _ = kinetic_energy(g, m, I)
        ts, kes, cts, cs = [0.0], [ke0], [], []
        total = 0.0
        done = 0
        while done < args.full_run:
            k = min(1000, args.full_run - done)
            total += g.step_timed(k)
            done += k
            ke, st = kinetic_energy(g, m, I)
            ts.append(done * wl["dt"])
            kes.append(ke)
            if done % 5000 == 0:
                cts.append(done * wl["dt"])
                cs.append(pair_contacts(st["x"] - np.asarray(wl["origin"]), r, box) / len(r))
        t = np.asarray(ts)
        ke = np.asarray(kes)
        sel = t >= 5000 * wl["dt"]
        A = np.vstack([np.ones(sel.sum()), t[sel]]).T
        coef, *_ = np.linalg.lstsq(A, ke[sel] ** -0.5, rcond=None)
        fit = (coef[0] + coef[1] * t[sel]) ** -2
        tau = float(coef[0] / coef[1]) if coef[1] > 0 else None
        out[nm] = {"ms": total, "particle_steps_per_s": len(r) * args.full_run / (total / 1e3),
                   "haff_tau_s": tau, "haff_fit_max_rel_dev": float(np.max(np.abs(fit / ke[sel] - 1))),
                   "ke_t0": ke0, "ke_end": float(ke[-1]), "contacts_per_particle": dict(zip(map(str, cts), cs))}
        curves[nm] = ke
        log(f"bench: full {args.full_run}-step run {nm}: {total:.0f} ms ({out[nm]['particle_steps_per_s']:.4g} "
            f"particle-steps/s), KE {ke0:.4g} -> {ke[-1]:.4g} J, Haff tau {tau} s, contacts/particle {cs}")
        del g
    rel = np.abs(curves["mixed"] / curves["fp64"] - 1)
    out["ke_mixed_vs_fp64_max_rel"] = float(rel.max())
    out["ke_samples"] = {"t_s": [float(x) for x in np.asarray(ts)], "fp64": [float(x) for x in curves["fp64"]],
                         "mixed": [float(x) for x in curves["mixed"]]}
    log(f"bench: kinetic energy mixed vs FP64 over the run: max relative difference {rel.max():.2e}")
    return out


def c1_leg(args):
    """C1: 1,000 spheres settling in a 20 x 20 x 60 mm box of walls, FP64,
    10,000 steps timed in one call; the CPU oracle on the same run, bounded
    to its first 2,000 steps."""
    from pacific_amd import FP64, workloads
    wl = workloads.settling_box(1000)
    g = build_gpu(wl, FP64)
    steps = args.c1_steps
    ms = g.step_timed(steps)
    vmax, vmean = g.velocity_max_mean()
    value = 1000 * steps / (ms / 1e3)
    out = {"c1_value": value, "c1_config": {"workload": "1000 spheres r 1 mm settling in a 20x20x60 mm box of "
                                                        "walls, FP64, Hooke kn 1e4 en 0.9 mu 0.5, dt 1e-5",
                                            "steps": steps, "ms": ms, "max_v_end": vmax, "mean_v_end": vmean}}
    log(f"bench: C1 settling 1000 spheres x {steps} FP64 steps: {ms:.0f} ms -> {value:.4g} particle-steps/s")
    if not args.no_cpu:
        v, t = cpu_baseline(wl, 2000)
        out["c1_cpu_baseline"] = {"value": v, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                                  "sample": f"the same run, its first 2000 steps ({t:.1f} s), FP64 oracle"}
        log(f"bench: C1 on the CPU oracle {v:.4g} particle-steps/s on 1 core")
    return out


def adjoint_leg(args):
    """Differentiable rollout: 20,000 spheres r~U[0.9,1.1] mm thrown onto a
    Box plane inclined at 25 degrees, 500 FP64 LeapFrog steps.  Ratio of the
    time of the checkpointed forward run (state kept every 50 steps) plus the
    reverse sweep (each segment recomputed from its checkpoint) for the
    gradient of the final centre-of-mass height, to the time of the plain
    500-step forward run; wall clock around synchronised calls, both sides."""
    from pacific_amd import FP64, workloads
    wl = workloads.inclined_plane()
    g = build_gpu(wl, FP64, skin=0.0)  # the adjoint replays the per-step neighbour search
    s0 = g.state()
    n, K = args.adjoint_steps, 50
    m, _, _ = g.masses()
    zN = np.zeros((g.n, 3))
    zN[:, 2] = m / m.sum()

    def reset():
        g.set_state(x=s0["x"], v=s0["v"], w=s0["w"], q=s0["q"], recompute=True)
        g.synchronize()

    best = {"plain": 1e30, "fwd": 1e30, "bwd": 1e30}
    for _ in range(3):
        reset()
        t0 = time.perf_counter()
        g.step(n)
        g.synchronize()
        best["plain"] = min(best["plain"], time.perf_counter() - t0)
        reset()
        t0 = time.perf_counter()
        g.adjoint_forward(n, K)
        t1 = time.perf_counter()
        grad = g.adjoint_backward(dJ_dx=zN)
        t2 = time.perf_counter()
        best["fwd"] = min(best["fwd"], t1 - t0)
        best["bwd"] = min(best["bwd"], t2 - t1)
    # the plain forward with the library's default contact search (Verlet
    # list), the fastest forward of this workload: the headline ratio
    gd = build_gpu(wl, FP64)
    sd = gd.state()
    best["default"] = 1e30
    for _ in range(3):
        gd.set_state(x=sd["x"], v=sd["v"], w=sd["w"], q=sd["q"], recompute=True)
        gd.synchronize()
        t0 = time.perf_counter()
        gd.step(n)
        gd.synchronize()
        best["default"] = min(best["default"], time.perf_counter() - t0)
    del gd
    ratio = (best["fwd"] + best["bwd"]) / best["default"]
    ratio_same = (best["fwd"] + best["bwd"]) / best["plain"]
    # the second objective of the workload, the final kinetic energy, from
    # the same forward run (outside the timed region)
    _, I, _ = g.masses()
    end = g.state()
    ke = float(0.5 * np.sum(m[:, None] * end["v"] ** 2) + 0.5 * np.sum(I[:, None] * end["w"] ** 2))
    gke = g.adjoint_backward(dJ_dv=m[:, None] * end["v"], dJ_dw=I[:, None] * end["w"])
    log(f"bench: adjoint {wl['name']} {n} steps K={K}: plain {best['default'] * 1e3:.1f} ms (default search), "
        f"{best['plain'] * 1e3:.1f} ms (per-step list), checkpointed forward {best['fwd'] * 1e3:.1f} ms, reverse "
        f"{best['bwd'] * 1e3:.1f} ms -> ratio {ratio:.3f} ({ratio_same:.3f} against the per-step-list forward)")
    return {"adjoint_ratio": ratio, "adjoint_ratio_vs_per_step_list_forward": ratio_same,
            "adjoint_config": {"workload": f"{wl['radius'].size} spheres r~U[0.9,1.1] mm onto a Box plane at 25 deg, "
                                           f"{n} steps dt=1e-5, FP64, checkpoints every {K}",
                               "plain_ms": best["default"] * 1e3,
                               "plain_per_step_list_ms": best["plain"] * 1e3,
                               "forward_checkpointed_ms": best["fwd"] * 1e3,
                               "reverse_ms": best["bwd"] * 1e3, "objective": "final centre-of-mass height",
                               "dJ_dparams": [float(x) for x in grad["params"]],
                               "dJ_dv0_norm": float(np.linalg.norm(grad["v0"])),
                               "kinetic_energy": {"value": ke, "dJ_dparams": [float(x) for x in gke["params"]],
                                                  "dJ_dv0_norm": float(np.linalg.norm(gke["v0"]))},
                               "params": "kn, en, muc of material pair (0, 0): sphere-sphere and sphere-plane"}}


def polyhedra_leg(args, prec):
    """200,000 polyhedra (thirds of cubes, tetrahedra, icosahedra) in a
    silo 0.1 m wide with a 12 mm orifice, released from rest: the lattice
    falls onto the floor plates and the bed builds up while the bodies above
    the orifice discharge (untimed fill), then the discharge is timed.  GJK
    narrow phase on crust-shrunk shapes, ParticleKinematics3D, basis
    composition."""
    from pacific_amd import FP64, workloads
    wl = workloads.silo(args.poly_n)
    col = {"gjk_sv": True} if args.poly_gjk_sv else None
    g = build_gpu(wl, prec, collision=col)
    t0 = time.perf_counter()
    done = 0
    if args.poly_state and os.path.exists(args.poly_state):
        # the filled silo of an earlier run of the same workload (profiling
        # runs skip the fill's dispatches)
        st = np.load(args.poly_state)
        g.set_basis(st["B"])
        g.set_state(x=st["x"], v=st["v"], w=st["w"], q=st["q"], recompute=True)
        g.set_time(float(st["t"]))
        done = args.poly_fill
        log(f"bench: silo state loaded from {args.poly_state}")
    while done < args.poly_fill:
        k = min(5000, args.poly_fill - done)
        g.step(k)
        g.synchronize()
        done += k
        log(f"bench: silo fill {done}/{args.poly_fill} steps ({time.perf_counter() - t0:.1f} s)")
    if args.poly_state and not os.path.exists(args.poly_state):
        st = g.state()
        np.savez(args.poly_state, x=st["x"], v=st["v"], w=st["w"], q=st["q"], B=g.basis(), t=g.time)
    g.step(args.poly_warmup)
    g.synchronize()
    ms = g.step_timed(args.poly_steps)
    value = args.poly_n * args.poly_steps / (ms / 1e3)
    g.record_pairs(True)
    g.compute_forces()
    pairs = g.pairs()
    g.record_pairs(False)
    cbar = len(pairs) / 2.0 / args.poly_n
    touching = len(np.unique(pairs[:, 0])) / args.poly_n if len(pairs) else 0.0
    prof = g.profile_steps(min(args.poly_steps, 100))
    force_ms = prof["force_ms"] / max(prof["n_force"], 1)
    narrow_ms = prof["narrow_ms"] / max(prof["n_narrow"], 1)
    detect_ms = prof["detect_ms"] / max(prof["n_detect"], 1)
    ferr = None
    if prec != FP64:
        g64 = build_gpu(wl, FP64, collision=col)
        ferr = force_error(g, g64, shapes=True)
        del g64
        log(f"bench: polyhedra mixed vs FP64 force error {ferr.get('max_rel_per_particle_sum', 0.0):.2e} "
            f"(per particle, relative to the sum of its contact-force magnitudes; "
            f"{ferr['particles_with_contacts']} particles in contact)")
    log(f"bench: polyhedra silo {args.poly_n}: {value:.4g} particle-steps/s, narrow phase {narrow_ms * 1e3:.1f} us, "
        f"force sweep {force_ms * 1e3:.1f} us, candidates {detect_ms * 1e3:.1f} us, contacts/particle {cbar:.3f}, "
        f"bodies in contact {touching:.3f}")
    return {"polyhedra_value": value,
            "polyhedra_config": {"workload": f"silo of {args.poly_n} polyhedra (cube 2 mm, tetrahedron and "
                                             "icosahedron R 1.2 mm), random orientations, released from rest, "
                                             "GJK, dt 5e-6", "fill_steps": args.poly_fill,
                                 "gjk": "GJK_SV" if args.poly_gjk_sv else "Johnson",
                                 "warmup": args.poly_warmup, "steps": args.poly_steps,
                                 "ms_per_step": ms / args.poly_steps, "narrow_phase_ms": narrow_ms,
                                 "force_sweep_ms": force_ms, "candidate_search_ms": detect_ms,
                                 "contacts_per_particle": cbar, "fraction_in_contact": touching,
                                 "mixed_force_error": ferr}}


def force_error(gm, g64, shapes=False):
    """Mixed-precision force error at one state: the FP64 system is put in
    the mixed system's state and both evaluate forces there.  Per particle
    with contacts, |F_mixed - F_fp64| relative to the sum of the magnitudes
    of its contact forces in the FP64 evaluation (sum |f_contact|): the
    largest and the mean over particles, and the 99.9th percentile; the
    same for torques relative to sum |f_contact| x the particle's largest
    lever (its circumscribed radius); particles without contacts must carry
    the FP64 force exactly (counted in `exact_without_contacts`)."""
    s = gm.state()
    if shapes:
        g64.set_basis(gm.basis())
    g64.set_state(x=s["x"], v=s["v"], w=s["w"], q=s["q"], recompute=False)
    fs = g64.force_sums()
    gm.compute_forces()
    a, b = gm.state(), g64.state()
    touch = fs > 0
    out = {"particles_with_contacts": int(touch.sum())}
    dF = np.linalg.norm(a["F"] - b["F"], axis=1)
    dT = np.linalg.norm(a["T"] - b["T"], axis=1)
    out["exact_without_contacts"] = bool(np.all(dF[~touch] == 0.0))
    if touch.any():
        eF = dF[touch] / fs[touch]
        lever = 1.2e-3 if not shapes else 1.74e-3
        eT = dT[touch] / (fs[touch] * lever)
        out["max_rel_per_particle_sum"] = float(eF.max())
        out["mean_rel_per_particle_sum"] = float(eF.mean())
        out["p999_rel_per_particle_sum"] = float(np.quantile(eF, 0.999))
        out["max_rel_torque_per_particle_sum"] = float(eT.max())
        out["mean_rel_torque_per_particle_sum"] = float(eT.mean())
    return out


def settling_force_error():
    """C1 (1,000 spheres settling in a box of walls) run to rest in FP64,
    then the mixed force error at that state."""
    from pacific_amd import FP64, MIXED, workloads
    wl = workloads.settling_box(1000)
    g64 = build_gpu(wl, FP64)
    g64.step(15000)
    gm = build_gpu(wl, MIXED)
    s = g64.state()
    gm.set_state(x=s["x"], v=s["v"], w=s["w"], q=s["q"], recompute=True)
    return force_error(gm, g64)


def read_traffic(name):
    """HBM bytes per launch of the force sweep from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f).get("bytes_per_launch")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="fp64",
                    help="precision of the headline value (the other one is measured beside it)")
    ap.add_argument("--relax", type=int, default=5000,
                    help="untimed steps out of the lattice start before the warmup (collisions established)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the second-precision leg and the force error")
    ap.add_argument("--full-run", type=int, default=20000,
                    help="steps of the full cooling run timed once per precision, with kinetic-energy and "
                         "coordination samples (0: skip)")
    ap.add_argument("--no-c1", action="store_true", help="skip the 1,000-sphere settling leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-ensemble-systems", type=int, default=128,
                    help="ensemble systems in the all-cores CPU leg (0: skip it)")
    ap.add_argument("--cpu-ensemble-steps", type=int, default=5000)
    ap.add_argument("--profile-steps", type=int, default=400)
    ap.add_argument("--nbr-cap", type=int, default=32,
                    help="neighbour candidates per particle of the per-step list (0: 27-cell sweep in the force kernel)")
    ap.add_argument("--skin", type=float, default=-1.0,
                    help="Verlet skin in m (< 0: a tenth of the cell size; 0: candidate list rebuilt every step)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--batch-steps", type=int, default=5000)
    ap.add_argument("--batch-chunk", type=int, default=100,
                    help="steps per captured hipGraph of the batch (0: direct launches, for tracers)")
    ap.add_argument("--batch-settle", type=int, default=15000,
                    help="untimed ensemble steps before the tuner (every layer has landed, the beds form)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward cost-ratio leg")
    ap.add_argument("--no-poly", action="store_true", help="skip the polyhedra (GJK) leg")
    ap.add_argument("--poly-n", type=int, default=200_000)
    ap.add_argument("--poly-steps", type=int, default=400)
    ap.add_argument("--poly-warmup", type=int, default=100)
    ap.add_argument("--poly-fill", type=int, default=20000, help="untimed steps of the silo fill from rest")
    ap.add_argument("--poly-state", default="", metavar="NPZ",
                    help="load the filled silo from this file if it exists, else save it there after the fill")
    ap.add_argument("--poly-gjk-sv", action="store_true",
                    help="the polyhedra leg with <CollisionDetection Type=\"GJK_SV\"> (default: Johnson's GJK)")
    ap.add_argument("--only-poly", action="store_true", help="run the polyhedra leg alone (profiling)")
    ap.add_argument("--adjoint-steps", type=int, default=500)
    ap.add_argument("--only-adjoint", action="store_true", help="run the adjoint leg alone (profiling)")
    ap.add_argument("--only-batch", action="store_true", help="run the batched-ensemble leg alone (profiling)")
    ap.add_argument("--c1-steps", type=int, default=10000)
    ap.add_argument("--only-c1", action="store_true", help="run the 1,000-sphere settling leg alone (profiling)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (counter calibration)")
    args = ap.parse_args()
    global NBR_CAP, SKIN
    NBR_CAP = args.nbr_cap
    SKIN = args.skin

    if args.calibrate_copy:
        # a plain copy of known size: the rocprofv3 FETCH_SIZE / WRITE_SIZE of
        # this kernel calibrates the counters used for the force sweep
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 8, "reps": 5}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs the DEM step on one GPU only (no sharding, no collectives); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.only_adjoint:
        print(json.dumps(adjoint_leg(args)), flush=True)
        return
    if args.only_batch:
        from pacific_amd import FP64
        print(json.dumps(batched_leg(args, FP64)), flush=True)
        return
    if args.only_c1:
        args.no_cpu = True
        print(json.dumps(c1_leg(args)), flush=True)
        return
    if args.only_poly:
        from pacific_amd import FP64, MIXED
        print(json.dumps(polyhedra_leg(args, MIXED if args.precision == "mixed" else FP64)), flush=True)
        return

    # the all-cores CPU leg of the ensemble first, while no process has
    # touched the GPU (its workers are forked)
    ens_cpu = None
    if not args.no_cpu and not args.no_batch and args.cpu_ensemble_systems > 0:
        cores = min(16, len(os.sched_getaffinity(0)))
        v, t = ensemble_cpu_baseline(args.cpu_ensemble_systems, args.cpu_ensemble_steps, cores)
        ens_cpu = {"value": v, "unit": "particle-steps/s", "cores": cores, "kind": "port",
                   "sample": f"the first {args.cpu_ensemble_systems} of the 512 ensemble systems, "
                             f"{args.cpu_ensemble_steps} steps each (the full settling run), one system per "
                             f"worker process ({t:.1f} s wall), FP64 oracle"}
        log(f"bench: CPU oracle on the ensemble {v:.4g} particle-steps/s on {cores} cores")

    from pacific_amd import FP64, MIXED, workloads
    wl = workloads.periodic_gas(args.n, seed=1)
    headline = FP64 if args.precision == "fp64" else MIXED
    other = MIXED if headline == FP64 else FP64
    name = {FP64: "fp64", MIXED: "mixed"}

    def gas_leg(prec):
        """The cooling gas after a fixed untimed relaxation out of the
        lattice start (collisions established), then W untimed and K timed
        steps."""
        log(f"bench: building {wl['name']} ({name[prec]}), relaxing {args.relax} steps untimed")
        g = build_gpu(wl, prec)
        g.step(args.relax)
        g.step(args.warmup)
        g.synchronize()
        t0 = time.perf_counter()
        ms = g.step_timed(args.steps)
        g.synchronize()
        wall = time.perf_counter() - t0
        v = args.n * args.steps / (ms / 1e3)
        cbar = contacts_per_particle(g)
        log(f"bench: {name[prec]}: {args.steps} steps in {ms:.2f} ms (wall {wall:.3f} s) -> {v:.4g} "
            f"particle-steps/s, contacts/particle {cbar:.4f} at the timed state")
        return g, ms, v, cbar

    g, ms, value, cbar = gas_leg(headline)
    prof = g.profile_steps(args.profile_steps)
    force_ms = prof["force_ms"] / max(prof["n_force"], 1)
    bin_ms = prof["bin_ms"] / max(prof["n_bin"], 1)
    detect_ms = prof["detect_ms"] / max(prof["n_detect"], 1) if prof["n_detect"] else 0.0
    b_alg = (B_ALG[name[headline]] + B_PAIR * cbar) * args.n
    achieved = b_alg / (force_ms / 1e3) / 1e9
    skin, builds = g.neighbor_stats()
    ngraphs, replays, why = g.graph_stats()
    log(f"bench: step graphs {ngraphs} ({replays} replays) {why}")
    log(f"bench: force sweep {force_ms * 1e3:.1f} us/launch, contact filter {detect_ms * 1e3:.1f} us/launch, "
        f"binning + list rebuild {bin_ms * 1e3:.1f} us/step, Verlet skin {skin:.3g} m, {builds} list builds in "
        f"{args.relax + args.warmup + args.steps + args.profile_steps} steps")

    extra = {}
    if not args.no_fp64:
        g2, ms2, v2, c2 = gas_leg(other)
        extra[f"{name[other]}_value"] = v2
        extra[f"{name[other]}_ms_per_step"] = ms2 / args.steps
        extra[f"{name[other]}_contacts_per_particle"] = c2
        gm, g64 = (g2, g) if other == MIXED else (g, g2)
        fe = {"gas": force_error(gm, g64)}
        del g2, gm, g64
        fe["settling_box_1000_at_rest"] = settling_force_error()
        extra["mixed_force_error"] = fe
        log(f"bench: mixed vs FP64 force error per particle / sum|f_contact|: gas max "
            f"{fe['gas'].get('max_rel_per_particle_sum', 0):.2e} mean {fe['gas'].get('mean_rel_per_particle_sum', 0):.2e}; "
            f"settled box max {fe['settling_box_1000_at_rest'].get('max_rel_per_particle_sum', 0):.2e}")
    del g
    if args.full_run > 0:
        extra["full_run"] = full_run_leg(args, wl)
    if not args.no_c1:
        extra.update(c1_leg(args))
    if not args.no_batch:
        extra.update(batched_leg(args, FP64))
    if not args.no_adjoint:
        extra.update(adjoint_leg(args))
    if not args.no_poly:
        extra.update(polyhedra_leg(args, MIXED))

    traffic = read_traffic(f"traffic_forces_{name[headline]}_{args.n}.json")
    out = {
        "metric": "particle-steps/s on 1 GPU, plain FP64 vs mixed/batched; adjoint/forward cost ratio",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-force/f64-state" if headline == MIXED else "f64",
        "data": "synthetic",
        "config": {"workload": f"periodic cooling granular gas, {args.n} polydisperse spheres r~U[0.8,1.2] mm, "
                               "phi=0.3, sigma_v=0.1 m/s, Hooke kn=1e4 en=0.8 mu=0.3, dt=1e-5, LeapFrog",
                   "particles": args.n, "precision": name[headline], "untimed_relaxation_steps": args.relax,
                   "contacts_per_particle": cbar, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"k_forces{'64' if headline == FP64 else ''}<fused>", "kernel_ms": force_ms,
                     "bytes_alg_per_launch": b_alg, "contacts_per_particle": cbar, "binning_ms": bin_ms,
                     "contact_filter_ms": detect_ms, "verlet_skin_m": skin, "verlet_builds": builds},
    }
    out.update(extra)
    if not args.no_cpu:
        cpu_value, cpu_s = cpu_baseline(wl, args.cpu_steps)
        out["cpu_baseline"] = {"value": cpu_value, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                               "sample": f"the same {args.n}-particle gas from its lattice start, {args.cpu_steps} "
                                         f"steps ({cpu_s:.1f} s), FP64 serial oracle restating the reference step",
                               "host": host_cpu()}
        log(f"bench: CPU oracle {cpu_value:.4g} particle-steps/s on 1 core")
        if ens_cpu is not None:
            out["batched_cpu_baseline"] = ens_cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
