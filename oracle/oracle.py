"""ctypes wrapper over the CPU oracle (oracle/grains_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  Never part of the product
path.  The library is built by ``make -C oracle`` (or __graft_entry__.build()).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")
_lib = None

# integrator ids (Scheme enum in grains_oracle.cpp)
LEAPFROG, FIRST_ORDER_EXPLICIT, SECOND_ORDER_EXPLICIT, ADAMS_BASHFORTH2 = 0, 1, 2, 3
ERRORS = {0: "OK", 1: "ContactError", 2: "MotionError", 3: "SimulationError", 4: "ConfigError"}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        vp = ctypes.c_void_p
        L.orc_new.restype = vp
        L.orc_new.argtypes = [dp, dp, ip, ctypes.c_double, dp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_free.argtypes = [vp]
        L.orc_set_contact_model.argtypes = [vp, ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 5
        L.orc_set_memory_contact_model.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.orc_add_sphere.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     dp, dp, dp, dp]
        L.orc_add_spheres.argtypes = [vp, ctypes.c_int, dp, dp, dp, ip, dp, dp, dp]
        L.orc_add_box.argtypes = [vp, dp, dp, dp, ctypes.c_double, ctypes.c_int]
        L.orc_init.argtypes = [vp]
        L.orc_set_collision.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_bv_overlap.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int, dp, dp]
        L.orc_step.argtypes = [vp, ctypes.c_int]
        L.orc_num_particles.argtypes = [vp]
        L.orc_num_active.argtypes = [vp]
        L.orc_time.argtypes = [vp]
        L.orc_time.restype = ctypes.c_double
        L.orc_get_state.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp]
        L.orc_get_mass.argtypes = [vp, dp, dp, dp]
        L.orc_get_force_sums.argtypes = [vp, dp]
        L.orc_get_term_scales.argtypes = [vp, dp]
        L.orc_get_obstacle_loads.argtypes = [vp, dp, dp]
        L.orc_get_pairs.argtypes = [vp, ip, ctypes.c_int]
        L.orc_get_grid.argtypes = [vp, ip, dp, dp]
        L.orc_get_cells.argtypes = [vp, ip]
        L.orc_get_stats.argtypes = [vp, dp]
        L.orc_set_state.argtypes = [vp, dp, dp, dp, dp, ctypes.c_int]
        L.orc_set_forces.argtypes = [vp, dp, dp]
        L.orc_move_only.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        L.orc_advance_only.argtypes = [vp, ctypes.c_double]
        L.orc_message.argtypes = [vp]
        L.orc_message.restype = ctypes.c_char_p
        L.orc_add_shape.argtypes = [vp, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int, dp, ctypes.c_int, ip, ip,
                                    ctypes.c_double]
        L.orc_add_particle.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp, dp, dp]
        L.orc_shape_info.argtypes = [vp, ctypes.c_int, dp]
        L.orc_gjk.argtypes = [vp, ctypes.c_int, dp, dp, ctypes.c_int, dp, dp, ctypes.c_int, dp, dp, ip]
        L.orc_gjk.restype = ctypes.c_double
        L.orc_get_basis.argtypes = [vp, dp]
        L.orc_set_basis.argtypes = [vp, dp]
        L.orc_closest_point.argtypes = [vp, ctypes.c_int, dp, dp, ctypes.c_int, dp, dp, dp]
        L.orc_gjk_stats.argtypes = [ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]
        L.orc_gjk_last.restype = ctypes.c_long
        _lib = L
    return _lib


def _d(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleError(RuntimeError):
    def __init__(self, code, ids=None):
        super().__init__(f"{ERRORS.get(code, code)} ids={ids}")
        self.code = code
        self.kind = ERRORS.get(code, str(code))
        self.ids = ids


class Oracle:
    """One serial Grains3D system (spheres + box obstacles)."""

    def __init__(self, origin, dmax, periodic=(0, 0, 0), cellsize_factor=1.0, gravity=(0.0, 0.0, -9.81),
                 fluid_density=0.0, dt=1e-5, scheme=LEAPFROG):
        L = lib()
        self._keep = []
        o = np.asarray(origin, np.float64)
        m = np.asarray(dmax, np.float64)
        g = np.asarray(gravity, np.float64)
        per = (ctypes.c_int * 3)(*[int(p) for p in periodic])
        self.h = L.orc_new(_d(o), _d(m), per, cellsize_factor, _d(g), fluid_density, dt, scheme)
        self.n = 0
        self.nobs = 0

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_free(self.h)
            self.h = None

    def contact_model(self, matA, matB, kn, en, muc, kr=0.0, etat=-1.0):
        lib().orc_set_contact_model(self.h, matA, matB, kn, en, etat, muc, kr)

    def collision(self, gjk_sv=False, tolerance=1e-10, acceleration=False, bounding_volume=0):
        """<CollisionDetectionAlgorithm>: GJK_SV or Johnson, relative
        tolerance, Nesterov acceleration, pre-test 0 none / 1 OBB / 2 OBC."""
        if lib().orc_set_collision(self.h, int(bool(gjk_sv)), float(tolerance), int(bool(acceleration)),
                                   int(bounding_volume)):
            raise ValueError("bad collision options")

    def bv_overlap(self, bv, shapeA, BA, oA, shapeB, BB, oB):
        """Bounding-volume pre-test (1 OBB, 2 OBC) of two placed shapes."""
        a = [np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1)) for x in (BA, oA, BB, oB)]
        return bool(lib().orc_bv_overlap(self.h, int(bv), int(shapeA), _d(a[0]), _d(a[1]), int(shapeB), _d(a[2]),
                                         _d(a[3])))

    def memory_contact_model(self, matA, matB, kn, en, muc, kt, etat=-1.0, mur=0.0, etarpf=0.0, epst=0.0):
        """HookeMemory (tangential spring kt, rolling friction mur)."""
        p = np.array([kn, en, etat, muc, kt, mur, etarpf, epst], np.float64)
        lib().orc_set_memory_contact_model(self.h, matA, matB, p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def add_spheres(self, radius, density, crust, material, x, v=None, w=None):
        x = np.asarray(x, np.float64).reshape(-1, 3)
        n = len(x)
        radius = np.broadcast_to(np.asarray(radius, np.float64), (n,))
        density = np.broadcast_to(np.asarray(density, np.float64), (n,))
        crust = np.broadcast_to(np.asarray(crust, np.float64), (n,))
        material = np.broadcast_to(np.asarray(material, np.int64), (n,))
        v = np.zeros((n, 3)) if v is None else np.asarray(v, np.float64).reshape(-1, 3)
        w = np.zeros((n, 3)) if w is None else np.asarray(w, np.float64).reshape(-1, 3)
        rad = np.ascontiguousarray(radius, np.float64)
        rho = np.ascontiguousarray(density, np.float64)
        cr = np.ascontiguousarray(crust, np.float64)
        mat = np.ascontiguousarray(material, np.int32)
        x, v, w = (np.ascontiguousarray(a, np.float64) for a in (x, v, w))
        lib().orc_add_spheres(self.h, n, _d(rad), _d(rho), _d(cr), mat.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              _d(x), _d(v), _d(w))
        self.n += n

    def add_box(self, L, centre, basis=None, crust=5e-5, material=0):
        b = None if basis is None else np.asarray(basis, np.float64).reshape(9)
        lib().orc_add_box(self.h, _d(np.asarray(L, np.float64)), _d(np.asarray(centre, np.float64)),
                          _d(b), crust, material)
        self.nobs += 1

    SPHERE, BOX, POLYHEDRON = 0, 1, 2

    def add_shape(self, kind, radius=0.0, L=None, verts=None, faces=None, crust=5e-5):
        """Convex shape class: kind 0 sphere(radius), 1 box(L = edge lengths),
        2 polyhedron(verts [nv,3], faces = list of vertex-index lists)."""
        Lv = np.zeros(3) if L is None else np.ascontiguousarray(L, np.float64)
        nv, nf = 0, 0
        vv = np.zeros(3)
        fs = np.zeros(1, np.int32)
        fi = np.zeros(1, np.int32)
        if kind == 2:
            vv = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
            nv, nf = len(vv), len(faces)
            fs = np.ascontiguousarray([len(f) for f in faces], np.int32)
            fi = np.ascontiguousarray([i for f in faces for i in f], np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        idx = lib().orc_add_shape(self.h, int(kind), float(radius), _d(Lv), nv, _d(vv), nf, fs.ctypes.data_as(ip),
                                  fi.ctypes.data_as(ip), float(crust))
        if idx < 0:
            raise OracleError(4, None)
        return idx

    def add_particles(self, shape, density, material, x, v=None, w=None, basis=None):
        x = np.asarray(x, np.float64).reshape(-1, 3)
        n = len(x)
        v = np.zeros((n, 3)) if v is None else np.asarray(v, np.float64).reshape(-1, 3)
        w = np.zeros((n, 3)) if w is None else np.asarray(w, np.float64).reshape(-1, 3)
        density = np.broadcast_to(np.asarray(density, np.float64), (n,))
        material = np.broadcast_to(np.asarray(material, np.int64), (n,))
        B = None if basis is None else np.asarray(basis, np.float64).reshape(n, 9)
        for i in range(n):
            lib().orc_add_particle(self.h, int(shape), float(density[i]), int(material[i]), _d(x[i]), _d(v[i]),
                                   _d(w[i]), None if B is None else _d(B[i]))
        self.n += n

    def shape_info(self, shape):
        out = np.zeros(14)
        lib().orc_shape_info(self.h, int(shape), _ptr(out))
        return {"volume": out[0], "circR": out[1], "inertia": out[2:8].copy(), "scaling": out[8:11].copy(),
                "bbox_extent": out[11:14].copy()}

    def gjk(self, shapeA, BA, oA, shapeB, BB, oB, crust=False):
        """GJK distance and world witness points between two placed shapes."""
        pa, pb = np.zeros(3), np.zeros(3)
        it = ctypes.c_int()
        d = lib().orc_gjk(self.h, int(shapeA), _d(np.asarray(BA, np.float64).reshape(9)),
                          _d(np.asarray(oA, np.float64)), int(shapeB), _d(np.asarray(BB, np.float64).reshape(9)),
                          _d(np.asarray(oB, np.float64)), 1 if crust else 0, _ptr(pa), _ptr(pb), ctypes.byref(it))
        return d, pa, pb, it.value

    def basis(self):
        B = np.zeros((self.n, 9))
        lib().orc_get_basis(self.h, _ptr(B))
        return B.reshape(self.n, 3, 3)

    def closest_point(self, shapeA, xA, BA, shapeB, xB, BB):
        out = np.zeros(8)
        lib().orc_closest_point(self.h, int(shapeA), _d(np.asarray(xA, np.float64)),
                                _d(np.asarray(BA, np.float64).reshape(9)), int(shapeB),
                                _d(np.asarray(xB, np.float64)), _d(np.asarray(BB, np.float64).reshape(9)), _ptr(out))
        return out[0:3].copy(), out[3:6].copy(), out[6], int(out[7])

    @staticmethod
    def gjk_last_iterations():
        """Iterations of the last GJK distance computation (either algorithm)."""
        return lib().orc_gjk_last()

    def set_basis(self, B):
        lib().orc_set_basis(self.h, _d(np.asarray(B, np.float64).reshape(self.n, 9)))

    def _check(self, code):
        if code:
            st = self.stats()
            raise OracleError(code, (int(st["err_id0"]), int(st["err_id1"])))

    def init(self):
        self._check(lib().orc_init(self.h))

    def step(self, n=1):
        self._check(lib().orc_step(self.h, n))

    @property
    def time(self):
        return lib().orc_time(self.h)

    def state(self):
        n = self.n
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "w", "wb", "F", "T")}
        out["q"] = np.zeros((n, 4))
        lib().orc_get_state(self.h, *[_ptr(out[k]) for k in ("x", "v", "w", "wb", "q", "F", "T")])
        return out

    def masses(self):
        n = self.n
        m, I, wgt = np.zeros(n), np.zeros(n), np.zeros((n, 3))
        lib().orc_get_mass(self.h, _ptr(m), _ptr(I), _ptr(wgt))
        return m, I, wgt

    def force_sums(self):
        """Per particle, the sum of the magnitudes of the contact forces of the
        last force evaluation (sum |f_contact|, the force-error normaliser)."""
        out = np.zeros(self.n)
        lib().orc_get_force_sums(self.h, _ptr(out))
        return out

    def term_scales(self):
        """Per particle, the sum over its memoryless contacts of kn |pen| +
        (|gamman| + |gammat|) |v_rel|: the magnitude of the terms its contact
        forces are made of (the scale of reduced-precision rounding)."""
        out = np.zeros(self.n)
        lib().orc_get_term_scales(self.h, _ptr(out))
        return out

    def obstacle_loads(self):
        F, T = np.zeros((self.nobs, 3)), np.zeros((self.nobs, 3))
        lib().orc_get_obstacle_loads(self.h, _ptr(F), _ptr(T))
        return F, T

    def pairs(self):
        """Directed contact records (receiver id, partner id) of the last force evaluation."""
        L = lib()
        n = L.orc_get_pairs(self.h, None, 0)
        buf = (ctypes.c_int * (2 * max(n, 1)))()
        L.orc_get_pairs(self.h, buf, n)
        return np.frombuffer(buf, dtype=np.int32)[: 2 * n].reshape(-1, 2).copy()

    def grid(self):
        dims = (ctypes.c_int * 3)()
        o, e = np.zeros(3), np.zeros(3)
        lib().orc_get_grid(self.h, dims, _ptr(o), _ptr(e))
        return tuple(dims), o, e

    def cells(self):
        buf = (ctypes.c_int * (3 * self.n))()
        lib().orc_get_cells(self.h, buf)
        return np.frombuffer(buf, dtype=np.int32).reshape(-1, 3).copy()

    def stats(self):
        out = np.zeros(9)
        lib().orc_get_stats(self.h, _ptr(out))
        return {"overlap_max": out[0], "overlap_mean": out[1], "recovered": out[2], "err_id0": out[3],
                "err_id1": out[4], "time_overlap_max": out[5], "gjk_iter_mean": out[6], "n_overlaps": out[7],
                "n_evaluations": out[8]}

    def set_state(self, x=None, v=None, w=None, q=None, recompute=True):
        keep = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (x, v, w, q)]
        self._check(lib().orc_set_state(self.h, *[_d(a) for a in keep], 1 if recompute else 0))

    def set_forces(self, F, T):
        F = np.ascontiguousarray(F, np.float64)
        T = np.ascontiguousarray(T, np.float64)
        lib().orc_set_forces(self.h, _d(F), _d(T))

    def move_only(self, dtv, dtx):
        self._check(lib().orc_move_only(self.h, dtv, dtx))

    def advance_only(self, dtv):
        lib().orc_advance_only(self.h, dtv)


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
