"""Synthetic DEM workloads with the shapes of the benchmark configurations.

Each builder returns a plain dict that both the HIP path (``GrainsGPU``) and
a checker can be filled from:

    domain: origin, max, periodic, cellsize_factor, gravity, dt
    particles: radius[n], density[n], crust[n], material[n], x[n,3], v[n,3], w[n,3]
    boxes: list of (L[3], centre[3], basis[9] or None, crust, material)
    models: list of (matA, matB, kn, en, muc, kr, etat)

``settling_box``  -- 1,000 (or n) spheres r = 1 mm, rho = 2500 kg/m^3 on a
                     jittered lattice in a 20 x 20 x 60 mm box of six thin
                     walls, gravity, kn = 1e4, en = 0.9, mu = 0.5, dt = 1e-5.
``periodic_gas``  -- n polydisperse spheres r ~ U[0.8, 1.2] mm in a triply
                     periodic cube at solid fraction 0.3, Gaussian velocities
                     (sigma 0.1 m/s), no gravity, kn = 1e4, en = 0.8,
                     mu = 0.3, dt = 1e-5 (homogeneous cooling granular gas).
"""
import numpy as np


def settling_box(n=1000, seed=12345, radius=1e-3, density=2500.0, crust=5e-5, kn=1e4, en=0.9, muc=0.5, kr=0.0,
                 dt=1e-5, box=(0.02, 0.02, 0.06), spacing=(2.5e-3, 2.5e-3, 3.75e-3), jitter=0.2e-3):
    bx, by, bz = box
    nx, ny, nz = int(round(bx / spacing[0])), int(round(by / spacing[1])), int(round(bz / spacing[2]))
    if n > nx * ny * nz:
        raise ValueError("too many particles for the lattice")
    # sites in z-major order (x fastest)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    sites = np.stack([(i.ravel() + 0.5) * spacing[0], (j.ravel() + 0.5) * spacing[1],
                      (k.ravel() + 0.5) * spacing[2]], axis=1)[:n]
    rng = np.random.default_rng(seed)
    x = sites + rng.uniform(-jitter, jitter, size=(n, 3))
    t = 0.5e-3  # wall thickness
    o = 0.5e-3  # overhang
    boxes = [
        ((bx + 2 * o, by + 2 * o, t), (bx / 2, by / 2, -t / 2)),
        ((bx + 2 * o, by + 2 * o, t), (bx / 2, by / 2, bz + t / 2)),
        ((t, by + 2 * o, bz + 2 * o), (-t / 2, by / 2, bz / 2)),
        ((t, by + 2 * o, bz + 2 * o), (bx + t / 2, by / 2, bz / 2)),
        ((bx + 2 * o, t, bz + 2 * o), (bx / 2, -t / 2, bz / 2)),
        ((bx + 2 * o, t, bz + 2 * o), (bx / 2, by + t / 2, bz / 2)),
    ]
    return {
        "name": f"settling_box_{n}",
        "origin": (0.0, 0.0, 0.0), "max": tuple(box), "periodic": (0, 0, 0), "cellsize_factor": 1.0,
        "gravity": (0.0, 0.0, -9.81), "dt": dt,
        "radius": np.full(n, radius), "density": np.full(n, density), "crust": np.full(n, crust),
        "material": np.zeros(n, np.int32), "x": x, "v": np.zeros((n, 3)), "w": np.zeros((n, 3)),
        "boxes": [(L, c, None, crust, 0) for L, c in boxes],
        "models": [(0, 0, kn, en, muc, kr, -1.0)],
    }


def periodic_gas(n=1_000_000, seed=1, phi=0.3, rmin=0.8e-3, rmax=1.2e-3, density=2500.0, crust=5e-5, sigma=0.1,
                 kn=1e4, en=0.8, muc=0.3, kr=0.0, dt=1e-5, jitter_frac=0.25):
    """Lattice placement: m^3 >= n sites, the first n used, jittered by a
    fraction of the smallest gap so that no two spheres overlap."""
    rng_r = np.random.default_rng(seed)
    radius = rng_r.uniform(rmin, rmax, size=n)
    vol = float(np.sum(4.0 / 3.0 * np.pi * radius ** 3))
    L = (vol / phi) ** (1.0 / 3.0)
    m = int(np.ceil(n ** (1.0 / 3.0)))
    a = L / m
    gap = a - 2 * rmax
    if gap <= 0:
        raise ValueError("solid fraction too high for a lattice start")
    rng_x = np.random.default_rng(seed + 1)
    idx = np.arange(n)
    i, j, k = idx % m, (idx // m) % m, idx // (m * m)
    x = np.stack([(i + 0.5) * a, (j + 0.5) * a, (k + 0.5) * a], axis=1)
    x += rng_x.uniform(-jitter_frac * gap / 2, jitter_frac * gap / 2, size=(n, 3))
    rng_v = np.random.default_rng(seed + 2)
    v = rng_v.normal(0.0, sigma, size=(n, 3))
    v -= v.mean(axis=0)
    return {
        "name": f"periodic_gas_{n}",
        "origin": (0.0, 0.0, 0.0), "max": (L, L, L), "periodic": (1, 1, 1), "cellsize_factor": 1.0,
        "gravity": (0.0, 0.0, 0.0), "dt": dt,
        "radius": radius, "density": np.full(n, density), "crust": np.full(n, crust),
        "material": np.zeros(n, np.int32), "x": x, "v": v, "w": np.zeros((n, 3)),
        "boxes": [], "models": [(0, 0, kn, en, muc, kr, -1.0)],
    }


def ensemble(nsys=512, seed=5, nmin=200, nmax=5000, radius=1e-3, density=2500.0, crust=5e-5, kn=1e4, dt=1e-5):
    """Independent settling boxes: N_i = round(exp(U[ln nmin, ln nmax])) spheres
    (seed), r = 1 mm, each on a C1-like lattice (spacing 2.5 x 2.5 x 3.75 mm,
    nx = ny = ceil((N/2)^(1/3)), nz = ceil(N / nx^2)) in its own walled box;
    per-system mu ~ U[0.1, 0.8] and en ~ U[0.5, 0.95] (seed + 1)."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), nsys))).astype(int)
    rp = np.random.default_rng(seed + 1)
    mus = rp.uniform(0.1, 0.8, nsys)
    ens = rp.uniform(0.5, 0.95, nsys)
    out = []
    for i, n in enumerate(sizes):
        nx = int(np.ceil((n / 2.0) ** (1.0 / 3.0)))
        nz = int(np.ceil(n / (nx * nx)))
        box = (nx * 2.5e-3, nx * 2.5e-3, nz * 3.75e-3)
        wl = settling_box(int(n), seed=seed * 1000 + i, radius=radius, density=density, crust=crust, kn=kn,
                          en=float(ens[i]), muc=float(mus[i]), dt=dt, box=box)
        wl["name"] = f"ensemble_{i}_{n}"
        out.append(wl)
    return out


def build(system, wl):
    """Fills a GrainsGPU (or any object with the same methods) from a workload."""
    for (ma, mb, kn, en, muc, kr, etat) in wl["models"]:
        system.contact_model(ma, mb, kn, en, muc, kr=kr, etat=etat)
    system.add_spheres(wl["radius"], wl["density"], wl["crust"], wl["material"], wl["x"], wl["v"], wl["w"])
    for (L, c, b, cr, mat) in wl["boxes"]:
        system.add_box(L, c, b, cr, mat)
    return system
