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


def build_gpu(wl, precision, skin=None):
    from pacific_amd import GrainsGPU, workloads
    g = GrainsGPU(wl["origin"], wl["max"], wl["periodic"], wl["cellsize_factor"], wl["gravity"], 0.0, wl["dt"],
                  0, precision)
    workloads.build(g, wl)
    g.neighbor_list(NBR_CAP, SKIN if skin is None else skin)
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


def batched_leg(args, prec):
    """512 independent settling boxes (sizes log-uniform in [200, 5000],
    per-system mu and en) in one persistent launch per hipGraph replay, the
    workgroup size picked by the autotuner."""
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
    best, tune = b.autotune(20)
    chunk = 100
    b.step(chunk, chunk=chunk)
    ms = b.step_timed(args.batch_steps, chunk=chunk)
    ntot = b.n_total
    value = ntot * args.batch_steps / (ms / 1e3)
    log(f"bench: batched ensemble 512 systems, {ntot} particles, workgroup {best}: {value:.4g} particle-steps/s")
    return {"batched_value": value, "batched_config": {"systems": 512, "particles": ntot, "steps": args.batch_steps,
                                                        "workgroup": best, "graph_chunk": chunk,
                                                        "autotune_ms": {str(k): v for k, v in tune.items()}}}


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
    ratio = (best["fwd"] + best["bwd"]) / best["plain"]
    # the second objective of the workload, the final kinetic energy, from
    # the same forward run (outside the timed region)
    _, I, _ = g.masses()
    end = g.state()
    ke = float(0.5 * np.sum(m[:, None] * end["v"] ** 2) + 0.5 * np.sum(I[:, None] * end["w"] ** 2))
    gke = g.adjoint_backward(dJ_dv=m[:, None] * end["v"], dJ_dw=I[:, None] * end["w"])
    log(f"bench: adjoint {wl['name']} {n} steps K={K}: plain {best['plain'] * 1e3:.1f} ms, checkpointed forward "
        f"{best['fwd'] * 1e3:.1f} ms, reverse {best['bwd'] * 1e3:.1f} ms -> ratio {ratio:.3f}")
    return {"adjoint_ratio": ratio,
            "adjoint_config": {"workload": f"{wl['radius'].size} spheres r~U[0.9,1.1] mm onto a Box plane at 25 deg, "
                                           f"{n} steps dt=1e-5, FP64, checkpoints every {K}",
                               "plain_ms": best["plain"] * 1e3, "forward_checkpointed_ms": best["fwd"] * 1e3,
                               "reverse_ms": best["bwd"] * 1e3, "objective": "final centre-of-mass height",
                               "dJ_dparams": [float(x) for x in grad["params"]],
                               "dJ_dv0_norm": float(np.linalg.norm(grad["v0"])),
                               "kinetic_energy": {"value": ke, "dJ_dparams": [float(x) for x in gke["params"]],
                                                  "dJ_dv0_norm": float(np.linalg.norm(gke["v0"]))},
                               "params": "kn, en, muc of material pair (0, 0): sphere-sphere and sphere-plane"}}


def polyhedra_leg(args, prec):
    """200,000 polyhedra (thirds of cubes, tetrahedra, icosahedra) in a
    silo, random orientations and velocities: GJK narrow phase on
    crust-shrunk shapes, ParticleKinematics3D, basis composition."""
    from pacific_amd import FP64, workloads
    wl = workloads.silo(args.poly_n)
    g = build_gpu(wl, prec)
    g.step(args.poly_warmup)
    g.synchronize()
    ms = g.step_timed(args.poly_steps)
    value = args.poly_n * args.poly_steps / (ms / 1e3)
    g.record_pairs(True)
    g.compute_forces()
    cbar = len(g.pairs()) / 2.0 / args.poly_n
    g.record_pairs(False)
    prof = g.profile_steps(min(args.poly_steps, 100))
    force_ms = prof["force_ms"] / max(prof["n_force"], 1)
    ferr = None
    if prec != FP64:
        g64 = build_gpu(wl, FP64)
        ferr = force_error(g, g64, shapes=True)
        del g64
        log(f"bench: polyhedra mixed vs FP64 force error {ferr['rel_to_max_F']:.2e} (relative to the largest force)")
    log(f"bench: polyhedra silo {args.poly_n}: {value:.4g} particle-steps/s, force sweep {force_ms * 1e3:.1f} us, "
        f"contacts/particle {cbar:.3f}")
    return {"polyhedra_value": value,
            "polyhedra_config": {"workload": f"silo of {args.poly_n} polyhedra (cube 2 mm, tetrahedron and "
                                             "icosahedron R 1.2 mm), random orientations, sigma_v 0.5 m/s, "
                                             "GJK, dt 5e-6", "steps": args.poly_steps,
                                 "ms_per_step": ms / args.poly_steps, "force_sweep_ms": force_ms,
                                 "contacts_per_particle": cbar, "mixed_force_error": ferr}}


def force_error(gm, g64, shapes=False):
    """Mixed-precision force error at one state: the FP64 system is put in
    the mixed system's state and both evaluate forces there.  Returns the
    largest force and torque deviations relative to the largest force /
    torque of the FP64 evaluation, and the largest per-particle deviation
    relative to that particle's contact-force scale (|F| with a floor of
    1e-6 of the largest)."""
    s = gm.state()
    if shapes:
        g64.set_basis(gm.basis())
    g64.set_state(x=s["x"], v=s["v"], w=s["w"], q=s["q"], recompute=True)
    gm.compute_forces()
    a, b = gm.state(), g64.state()
    out = {}
    for k in ("F", "T"):
        d = np.linalg.norm(a[k] - b[k], axis=1)
        m = np.linalg.norm(b[k], axis=1)
        top = float(m.max()) if m.size else 0.0
        out[f"max_abs_{k}"] = float(d.max()) if d.size else 0.0
        out[f"rel_to_max_{k}"] = float(d.max() / top) if top > 0 else 0.0
        out[f"max_rel_per_particle_{k}"] = float(np.max(d / np.maximum(m, 1e-6 * top))) if top > 0 else 0.0
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
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
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
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward cost-ratio leg")
    ap.add_argument("--no-poly", action="store_true", help="skip the polyhedra (GJK) leg")
    ap.add_argument("--poly-n", type=int, default=200_000)
    ap.add_argument("--poly-steps", type=int, default=400)
    ap.add_argument("--poly-warmup", type=int, default=100)
    ap.add_argument("--only-poly", action="store_true", help="run the polyhedra leg alone (profiling)")
    ap.add_argument("--adjoint-steps", type=int, default=500)
    ap.add_argument("--only-adjoint", action="store_true", help="run the adjoint leg alone (profiling)")
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
                             f"{args.cpu_ensemble_steps} steps each, one system per worker process "
                             f"({t:.1f} s wall), FP64 oracle"}
        log(f"bench: CPU oracle on the ensemble {v:.4g} particle-steps/s on {cores} cores")

    from pacific_amd import FP64, MIXED, workloads
    wl = workloads.periodic_gas(args.n, seed=1)
    prec = MIXED if args.precision == "mixed" else FP64
    log(f"bench: building {wl['name']} ({args.precision})")
    g = build_gpu(wl, prec)
    g.step(args.warmup)
    g.synchronize()
    t0 = time.perf_counter()
    ms = g.step_timed(args.steps)
    g.synchronize()
    wall = time.perf_counter() - t0
    value = args.n * args.steps / (ms / 1e3)
    log(f"bench: {args.steps} steps in {ms:.1f} ms (wall {wall:.3f} s) -> {value:.4g} particle-steps/s")

    cbar = contacts_per_particle(g)
    prof = g.profile_steps(args.profile_steps)
    force_ms = prof["force_ms"] / max(prof["n_force"], 1)
    bin_ms = prof["bin_ms"] / max(prof["n_bin"], 1)
    detect_ms = prof["detect_ms"] / max(prof["n_detect"], 1) if prof["n_detect"] else 0.0
    b_alg = (B_ALG[args.precision] + B_PAIR * cbar) * args.n
    achieved = b_alg / (force_ms / 1e3) / 1e9
    skin, builds = g.neighbor_stats()
    ngraphs, replays, why = g.graph_stats()
    log(f"bench: step graphs {ngraphs} ({replays} replays) {why}")
    log(f"bench: force sweep {force_ms * 1e3:.1f} us/launch, neighbour search {detect_ms * 1e3:.1f} us/launch, "
        f"binning {bin_ms * 1e3:.1f} us/step, contacts/particle {cbar:.4f}, Verlet skin {skin:.3g} m, "
        f"{builds} list builds in {args.warmup + args.steps + args.profile_steps} steps")

    extra = {}
    if not args.no_fp64 and prec != FP64:
        g64 = build_gpu(wl, FP64)
        g64.step(args.warmup)
        ms64 = g64.step_timed(max(args.steps // 2, 1))
        extra["fp64_value"] = args.n * max(args.steps // 2, 1) / (ms64 / 1e3)
        extra["fp64_ms_per_step"] = ms64 / max(args.steps // 2, 1)
        fe = {"gas": force_error(g, g64)}
        del g64
        fe["settling_box_1000_at_rest"] = settling_force_error()
        extra["mixed_force_error"] = fe
        log(f"bench: mixed vs FP64 force error: gas {fe['gas']['rel_to_max_F']:.2e}, settled box "
            f"{fe['settling_box_1000_at_rest']['rel_to_max_F']:.2e} (relative to the largest force)")
        log(f"bench: FP64 path {extra['fp64_value']:.4g} particle-steps/s")

    if not args.no_batch:
        extra.update(batched_leg(args, prec))
    if not args.no_adjoint:
        extra.update(adjoint_leg(args))
    if not args.no_poly:
        extra.update(polyhedra_leg(args, prec))

    traffic = read_traffic(f"traffic_forces_{args.precision}_{args.n}.json")
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
        "dtype": "f32-force/f64-state" if prec == MIXED else "f64",
        "data": "synthetic",
        "config": {"workload": f"periodic cooling granular gas, {args.n} polydisperse spheres r~U[0.8,1.2] mm, "
                               "phi=0.3, sigma_v=0.1 m/s, Hooke kn=1e4 en=0.8 mu=0.3, dt=1e-5, LeapFrog",
                   "particles": args.n, "precision": args.precision, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"k_forces<fused,{args.precision}>", "kernel_ms": force_ms,
                     "bytes_alg_per_launch": b_alg, "contacts_per_particle": cbar, "binning_ms": bin_ms,
                     "neighbour_search_ms": detect_ms, "verlet_skin_m": skin, "verlet_builds": builds},
    }
    out.update(extra)
    if not args.no_cpu:
        cpu_value, cpu_s = cpu_baseline(wl, args.cpu_steps)
        out["cpu_baseline"] = {"value": cpu_value, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                               "sample": f"the same {args.n}-particle gas, {args.cpu_steps} steps "
                                         f"({cpu_s:.1f} s), FP64 serial oracle restating the reference step",
                               "host": host_cpu()}
        log(f"bench: CPU oracle {cpu_value:.4g} particle-steps/s on 1 core")
        if ens_cpu is not None:
            out["batched_cpu_baseline"] = ens_cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
