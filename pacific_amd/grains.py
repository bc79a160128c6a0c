"""Python binding of libgrains_gpu.so (the C ABI in include/grains_gpu.h).

Mirrors the Grains3D objects that set up a DEM system -- domain, contact
models, spheres, box obstacles, time integration -- and exposes the state as
NumPy arrays in particle-id order.  The binding is ctypes over the C ABI: no
torch types cross the boundary, and there is no CPU fallback -- if the HIP
library is missing or no GPU is present, construction raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgrains_gpu.so")

LEAPFROG, FIRST_ORDER_EXPLICIT, SECOND_ORDER_EXPLICIT, ADAMS_BASHFORTH2 = 0, 1, 2, 3
FP64, MIXED = 0, 1
SCHEMES = {"SecondOrderLeapFrog": LEAPFROG, "FirstOrderExplicit": FIRST_ORDER_EXPLICIT,
           "SecondOrderExplicit": SECOND_ORDER_EXPLICIT, "SecondOrderAdamsBashforth": ADAMS_BASHFORTH2}
ERRORS = {0: "OK", 1: "ContactError", 2: "MotionError", 3: "SimulationError", 4: "ConfigError",
          5: "DeviceError", 6: "ArgumentError"}

_lib = None


class Config(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_double * 3), ("max", ctypes.c_double * 3), ("periodic", ctypes.c_int * 3),
                ("cell_size_factor", ctypes.c_double), ("gravity", ctypes.c_double * 3),
                ("fluid_density", ctypes.c_double), ("dt", ctypes.c_double), ("t_start", ctypes.c_double),
                ("scheme", ctypes.c_int), ("precision", ctypes.c_int), ("device", ctypes.c_int)]


EXPORTS = [
    "grains_gpu_create", "grains_gpu_create_from_xml", "grains_gpu_destroy", "grains_gpu_set_contact_model",
    "grains_gpu_add_spheres", "grains_gpu_add_box", "grains_gpu_initialize", "grains_gpu_step",
    "grains_gpu_step_timed", "grains_gpu_profile_steps", "grains_gpu_synchronize", "grains_gpu_compute_forces", "grains_gpu_get_state",
    "grains_gpu_set_state", "grains_gpu_get_masses", "grains_gpu_set_record_pairs", "grains_gpu_get_pairs",
    "grains_gpu_num_particles", "grains_gpu_time", "grains_gpu_set_time", "grains_gpu_grid",
    "grains_gpu_error_ids", "grains_gpu_recoveries", "grains_gpu_last_error",
    "grains_gpu_batch_create", "grains_gpu_batch_destroy", "grains_gpu_batch_set_block", "grains_gpu_batch_step",
    "grains_gpu_batch_step_timed", "grains_gpu_batch_autotune", "grains_gpu_batch_num_particles",
    "grains_gpu_batch_time", "grains_gpu_batch_get_state", "grains_gpu_batch_last_error",
    "grains_gpu_adjoint_forward", "grains_gpu_adjoint_backward", "grains_contact_eval",
    "grains_gpu_contact_counts", "grains_gpu_add_shape", "grains_gpu_add_particles", "grains_gpu_shape_info",
    "grains_gpu_get_basis", "grains_gpu_set_basis", "grains_gpu_create_coupled", "grains_gpu_fluid_step",
    "grains_gpu_to_fluid", "grains_gpu_set_velocities", "grains_gpu_count_rigid_bodies",
    "grains_gpu_time_interval", "grains_gpu_closest_point", "grains_gpu_set_neighbor_list",
    "grains_gpu_neighbor_stats", "grains_gpu_set_contact_stats", "grains_gpu_contact_stats",
    "grains_gpu_set_memory_contact_model", "grains_gpu_set_particle_class", "grains_gpu_velocity_max_mean",
    "grains_gpu_coordination_numbers", "grains_gpu_write_rawdata", "grains_gpu_set_outputs", "grains_gpu_run",
    "grains_gpu_graph_stats", "grains_gpu_force_sums", "grains_gpu_set_hydro_ft",
    "grains_gpu_coordination_at_state", "grains_gpu_post_processing", "grains_gpu_save_reload",
    "grains_gpu_load_reload", "grains_gpu_reload_same", "grains_gpu_force_sums_noalloc",
    "grains_gpu_set_collision", "grains_gpu_parse_xml", "grains_gpu_construction_state", "grains_gpu_try_insert",
    "grains_gpu_pending_insertions",
]


def lib():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build it with `python -m pacific_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        ip, sz = ctypes.POINTER(ctypes.c_int), ctypes.c_size_t
        L.grains_gpu_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(vp)]
        if hasattr(L, "grains_gpu_create_from_xml"):
            L.grains_gpu_create_from_xml.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(vp)]
        L.grains_gpu_destroy.argtypes = [vp]
        L.grains_gpu_destroy.restype = None
        L.grains_gpu_set_contact_model.argtypes = [vp, ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 5
        L.grains_gpu_add_spheres.argtypes = [vp, sz, dp, dp, dp, ip, dp, dp, dp]
        L.grains_gpu_add_box.argtypes = [vp, dp, dp, dp, ctypes.c_double, ctypes.c_int]
        L.grains_gpu_initialize.argtypes = [vp]
        L.grains_gpu_step.argtypes = [vp, ctypes.c_int]
        L.grains_gpu_step_timed.argtypes = [vp, ctypes.c_int, dp]
        L.grains_gpu_profile_steps.argtypes = [vp, ctypes.c_int, dp]
        L.grains_gpu_synchronize.argtypes = [vp]
        L.grains_gpu_compute_forces.argtypes = [vp]
        L.grains_gpu_force_sums.argtypes = [vp, dp]
        L.grains_gpu_force_sums_noalloc.argtypes = [vp, vp]
        L.grains_gpu_set_collision.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.grains_gpu_get_state.argtypes = [vp] + [dp] * 7
        L.grains_gpu_set_state.argtypes = [vp, dp, dp, dp, dp, ctypes.c_int]
        L.grains_gpu_get_masses.argtypes = [vp, dp, dp, dp]
        L.grains_gpu_set_record_pairs.argtypes = [vp, ctypes.c_int]
        L.grains_gpu_set_neighbor_list.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.grains_gpu_neighbor_stats.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_long)]
        L.grains_gpu_set_contact_stats.argtypes = [vp, ctypes.c_int]
        L.grains_gpu_set_memory_contact_model.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.grains_gpu_contact_stats.argtypes = [vp, dp]
        L.grains_gpu_set_particle_class.argtypes = [vp, ctypes.c_int]
        L.grains_gpu_velocity_max_mean.argtypes = [vp, dp, dp]
        L.grains_gpu_coordination_numbers.argtypes = [vp, ip]
        L.grains_gpu_coordination_at_state.argtypes = [vp, ip]
        L.grains_gpu_post_processing.argtypes = [vp, ctypes.c_int]
        L.grains_gpu_save_reload.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
        L.grains_gpu_load_reload.argtypes = [vp, ctypes.c_char_p]
        L.grains_gpu_reload_same.argtypes = [vp]
        L.grains_gpu_write_rawdata.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ip]
        L.grains_gpu_set_outputs.argtypes = [vp, dp, sz, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.grains_gpu_run.argtypes = [vp, ctypes.c_int]
        L.grains_gpu_construction_state.argtypes = [vp, dp, dp, dp, dp]
        L.grains_gpu_try_insert.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.grains_gpu_pending_insertions.argtypes = [vp, ctypes.POINTER(ctypes.c_long)]
        L.grains_gpu_graph_stats.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_long), ctypes.c_char_p, sz]
        L.grains_gpu_get_pairs.argtypes = [vp, ip, sz, ctypes.POINTER(sz)]
        L.grains_gpu_num_particles.argtypes = [vp]
        L.grains_gpu_num_particles.restype = sz
        L.grains_gpu_time.argtypes = [vp]
        L.grains_gpu_time.restype = ctypes.c_double
        L.grains_gpu_set_time.argtypes = [vp, ctypes.c_double]
        L.grains_gpu_set_time.restype = None
        L.grains_gpu_grid.argtypes = [vp, ip, dp]
        L.grains_gpu_error_ids.argtypes = [vp, ip, ip]
        L.grains_gpu_recoveries.argtypes = [vp, ctypes.POINTER(ctypes.c_long)]
        L.grains_gpu_last_error.argtypes = [vp]
        L.grains_gpu_last_error.restype = ctypes.c_char_p
        L.grains_gpu_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
        L.grains_gpu_batch_destroy.argtypes = [vp]
        L.grains_gpu_batch_destroy.restype = None
        L.grains_gpu_batch_set_block.argtypes = [vp, ctypes.c_int]
        L.grains_gpu_batch_step.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.grains_gpu_batch_step_timed.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.grains_gpu_batch_autotune.argtypes = [vp, ctypes.c_int, ip, dp]
        L.grains_gpu_batch_num_particles.argtypes = [vp, ctypes.c_int]
        L.grains_gpu_batch_num_particles.restype = sz
        L.grains_gpu_batch_time.argtypes = [vp]
        L.grains_gpu_batch_time.restype = ctypes.c_double
        L.grains_gpu_batch_get_state.argtypes = [vp, ctypes.c_int] + [dp] * 6
        L.grains_gpu_batch_last_error.argtypes = [vp]
        L.grains_gpu_batch_last_error.restype = ctypes.c_char_p
        L.grains_gpu_adjoint_forward.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.grains_gpu_adjoint_backward.argtypes = [vp, ctypes.c_int, ctypes.c_int] + [dp] * 7
        L.grains_contact_eval.argtypes = [ctypes.c_int] + [dp] * 6
        L.grains_gpu_contact_counts.argtypes = [vp, ctypes.POINTER(sz), ctypes.POINTER(sz)]
        L.grains_gpu_add_shape.argtypes = [vp, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int, dp, ctypes.c_int,
                                           ip, ip, ctypes.c_double, ip]
        L.grains_gpu_add_particles.argtypes = [vp, ctypes.c_int, sz, dp, ip, dp, dp, dp, dp]
        L.grains_gpu_shape_info.argtypes = [vp, ctypes.c_int, dp]
        L.grains_gpu_get_basis.argtypes = [vp, dp]
        L.grains_gpu_set_basis.argtypes = [vp, dp]
        L.grains_gpu_closest_point.argtypes = [vp, ctypes.c_int, dp, dp, ctypes.c_int, dp, dp, dp]
        L.grains_gpu_create_coupled.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_int, ctypes.POINTER(vp)]
        L.grains_gpu_fluid_step.argtypes = [vp, ctypes.c_double, ip]
        L.grains_gpu_to_fluid.argtypes = [vp, ctypes.POINTER(ctypes.c_void_p), ip]
        L.grains_gpu_set_velocities.argtypes = [vp, dp, sz]
        L.grains_gpu_set_hydro_ft.argtypes = [vp, dp, sz]
        L.grains_gpu_count_rigid_bodies.argtypes = [vp, ctypes.POINTER(sz), ctypes.POINTER(sz)]
        _lib = L
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _libc_free(p):
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    libc.free(p)


def _c64(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def contact_eval(kind, body0, body1, model, cot=None):
    """One Hooke contact through the host copy of the adjoint's derivative code
    (``grains_contact_eval``): returns (in_contact, loads[12], grad[12] or None).
    Needs the library but no device."""
    b0, b1, md = _c64(body0), _c64(body1), _c64(model)
    ct = _c64(cot)
    loads, grad = np.zeros(12), (np.zeros(12) if cot is not None else None)
    r = lib().grains_contact_eval(int(kind), _d(b0), _d(b1), _d(md), _d(ct), _d(loads), _d(grad))
    if r < 0:
        raise GrainsError(-r if -r in ERRORS else r, "grains_contact_eval: bad arguments")
    return bool(r), loads, grad


class GrainsError(RuntimeError):
    def __init__(self, code, msg, ids=(0, 0)):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code
        self.kind = ERRORS.get(code, str(code))
        self.ids = ids


class GrainsGPU:
    """One DEM system on one GPU (spheres + static box obstacles)."""

    def __init__(self, origin, dmax, periodic=(0, 0, 0), cellsize_factor=1.0, gravity=(0.0, 0.0, -9.81),
                 fluid_density=0.0, dt=1e-5, scheme=LEAPFROG, precision=FP64, t_start=0.0, device=0, _handle=None):
        L = lib()
        self.h = ctypes.c_void_p()
        self.n = 0
        self.nobs = 0
        if _handle is not None:
            self.h = _handle
            self.n = int(L.grains_gpu_num_particles(self.h))
            return
        cfg = Config()
        for a in range(3):
            cfg.origin[a] = origin[a]
            cfg.max[a] = dmax[a]
            cfg.periodic[a] = int(periodic[a])
            cfg.gravity[a] = gravity[a]
        cfg.cell_size_factor = cellsize_factor
        cfg.fluid_density = fluid_density
        cfg.dt = dt
        cfg.t_start = t_start
        cfg.scheme = scheme
        cfg.precision = precision
        cfg.device = device
        self._check(L.grains_gpu_create(ctypes.byref(cfg), ctypes.byref(self.h)))

    @classmethod
    def coupled(cls, path, fluid_rho, fluid_corrected=True):
        """System of a fluid-coupled input file (grains_gpu_create_coupled)."""
        L = lib()
        h = ctypes.c_void_p()
        code = L.grains_gpu_create_coupled(os.fsencode(path), float(fluid_rho), 1 if fluid_corrected else 0,
                                           ctypes.byref(h))
        if code:
            msg = L.grains_gpu_last_error(h).decode() if h.value else "cannot build the system"
            if h.value:
                L.grains_gpu_destroy(h)
            raise GrainsError(code, msg)
        return cls(None, None, _handle=h)

    def fluid_step(self, dt_fluid):
        n = ctypes.c_int()
        self._check(lib().grains_gpu_fluid_step(self.h, float(dt_fluid), ctypes.byref(n)))
        return n.value

    def to_fluid(self):
        """GrainsToFluid text (tokens joined by single spaces)."""
        L = lib()
        buf, size = ctypes.c_void_p(), ctypes.c_int()
        self._check(L.grains_gpu_to_fluid(self.h, ctypes.byref(buf), ctypes.byref(size)))
        try:
            return ctypes.string_at(buf.value).decode()
        finally:
            _libc_free(buf)

    def set_velocities(self, vw):
        vw = _c64(vw, (-1, 6))
        self._check(lib().grains_gpu_set_velocities(self.h, _d(vw), len(vw)))

    def set_hydro_ft(self, ft):
        """Fluid force and torque on each particle, rows (Fx, Fy, Fz, Tx, Ty, Tz)
        by id - 1 (GrainsCoupledWithFluid::updateParticlesHydroFT)."""
        ft = _c64(ft, (-1, 6))
        self._check(lib().grains_gpu_set_hydro_ft(self.h, _d(ft), len(ft)))

    def save_reload(self, root, hybrid=False):
        """__Format2024__ reload files <root>A|B, <root>A|B_Particles[.bin],
        <root>_MatContact.xml, <root>_RFTable.txt (Grains::saveReload)."""
        self._check(lib().grains_gpu_save_reload(self.h, os.fsencode(str(root)), 1 if hybrid else 0))

    def load_reload(self, path):
        """State, torsor, contact memory and time from a reload file."""
        self._check(lib().grains_gpu_load_reload(self.h, os.fsencode(str(path))))

    def count_rigid_bodies(self):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        self._check(lib().grains_gpu_count_rigid_bodies(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    @classmethod
    def from_xml(cls, path, precision=FP64, init=True):
        """A system built from a Grains3D input file; init=False stops before
        initialisation (no device needed): the construction data only."""
        h = ctypes.c_void_p()
        make = lib().grains_gpu_create_from_xml if init else lib().grains_gpu_parse_xml
        code = make(os.fsencode(path), precision, ctypes.byref(h))
        if code:
            msg = lib().grains_gpu_last_error(h).decode() if h.value else "create failed"
            if h.value:
                lib().grains_gpu_destroy(h)
            raise GrainsError(code, msg)
        return cls(None, None, _handle=h)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value and _lib is not None:
            _lib.grains_gpu_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        self.close()

    def _check(self, code):
        if code:
            L = lib()
            msg = L.grains_gpu_last_error(self.h).decode() if self.h.value else ""
            a, b = ctypes.c_int(), ctypes.c_int()
            if self.h.value:
                L.grains_gpu_error_ids(self.h, ctypes.byref(a), ctypes.byref(b))
            raise GrainsError(code, msg, (a.value, b.value))

    def contact_model(self, matA, matB, kn, en, muc, kr=0.0, etat=-1.0):
        self._check(lib().grains_gpu_set_contact_model(self.h, matA, matB, kn, en, etat, muc, kr))

    def collision(self, gjk_sv=False, tolerance=1e-10, acceleration=False, bounding_volume=0):
        """<CollisionDetectionAlgorithm>: GJK_SV or Johnson, relative
        tolerance, Nesterov acceleration, pre-test 0 none / 1 OBB / 2 OBC
        (before init)."""
        self._check(lib().grains_gpu_set_collision(self.h, int(bool(gjk_sv)), float(tolerance),
                                                   int(bool(acceleration)), int(bounding_volume)))

    def memory_contact_model(self, matA, matB, kn, en, muc, kt, etat=-1.0, mur=0.0, etarpf=0.0, epst=0.0):
        """HookeMemory contact model (tangential spring kt, rolling friction mur)."""
        p = np.array([kn, en, etat, muc, kt, mur, etarpf, epst], np.float64)
        self._check(lib().grains_gpu_set_memory_contact_model(self.h, int(matA), int(matB), _d(p)))

    def add_spheres(self, radius, density, crust, material, x, v=None, w=None):
        x = _c64(x, (-1, 3))
        n = len(x)
        rad = _c64(np.broadcast_to(radius, (n,)))
        rho = _c64(np.broadcast_to(density, (n,)))
        cr = _c64(np.broadcast_to(crust, (n,)))
        mat = np.ascontiguousarray(np.broadcast_to(material, (n,)), dtype=np.int32)
        v = _c64(v, (-1, 3))
        w = _c64(w, (-1, 3))
        self._check(lib().grains_gpu_add_spheres(self.h, n, _d(rad), _d(rho), _d(cr),
                                                 mat.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _d(x), _d(v),
                                                 _d(w)))
        self.n += n

    SPHERE, BOX, POLYHEDRON = 0, 1, 2

    def add_shape(self, kind, radius=0.0, L=None, verts=None, faces=None, crust=5e-5):
        """Convex shape class: kind 0 sphere(radius), 1 box(L = edge lengths),
        2 polyhedron(verts [nv, 3], faces = list of vertex-index lists).
        Returns the class index."""
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
        out = ctypes.c_int()
        self._check(lib().grains_gpu_add_shape(self.h, int(kind), float(radius), _d(Lv), nv, _d(vv), nf,
                                               fs.ctypes.data_as(ip), fi.ctypes.data_as(ip), float(crust),
                                               ctypes.byref(out)))
        return out.value

    def add_particles(self, shape, density, material, x, v=None, w=None, basis=None):
        x = _c64(x, (-1, 3))
        n = len(x)
        rho = _c64(np.broadcast_to(np.asarray(density, np.float64), (n,)))
        mat = np.ascontiguousarray(np.broadcast_to(np.asarray(material), (n,)), np.int32)
        v = None if v is None else _c64(v, (n, 3))
        w = None if w is None else _c64(w, (n, 3))
        B = None if basis is None else _c64(basis, (n, 9))
        self._check(lib().grains_gpu_add_particles(self.h, int(shape), n, _d(rho),
                                                   mat.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _d(x), _d(v),
                                                   _d(w), _d(B)))
        self.n += n

    def shape_info(self, shape):
        out = np.zeros(14)
        self._check(lib().grains_gpu_shape_info(self.h, int(shape), _d(out)))
        return {"volume": out[0], "circR": out[1], "inertia": out[2:8].copy(), "scaling": out[8:11].copy(),
                "bbox_extent": out[11:14].copy()}

    def closest_point(self, shapeA, xA, BA, shapeB, xB, BB):
        """Host evaluation of the device narrow phase: (P, ov, d, status)."""
        out = np.zeros(8)
        args = [_c64(a).reshape(-1) for a in (xA, BA, xB, BB)]
        self._check(lib().grains_gpu_closest_point(self.h, int(shapeA), _d(args[0]), _d(args[1]), int(shapeB),
                                                   _d(args[2]), _d(args[3]), _d(out)))
        return out[0:3].copy(), out[3:6].copy(), out[6], int(out[7])

    def basis(self):
        B = np.zeros((self.n, 9))
        self._check(lib().grains_gpu_get_basis(self.h, _d(B)))
        return B.reshape(self.n, 3, 3)

    def set_basis(self, B):
        B = _c64(B, (self.n, 9))
        self._check(lib().grains_gpu_set_basis(self.h, _d(B)))

    def add_box(self, L, centre, basis=None, crust=5e-5, material=0):
        Lv, c = _c64(L), _c64(centre)
        b = _c64(basis, (9,))
        self._check(lib().grains_gpu_add_box(self.h, _d(Lv), _d(c), _d(b), crust, material))
        self.nobs += 1

    def init(self):
        self._check(lib().grains_gpu_initialize(self.h))

    def step(self, n=1):
        self._check(lib().grains_gpu_step(self.h, int(n)))

    def step_timed(self, n):
        ms = ctypes.c_double()
        self._check(lib().grains_gpu_step_timed(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def profile_steps(self, n):
        """Runs n steps with HIP events around each kernel group; returns
        dict(total_ms, force_ms, n_force, bin_ms, n_bin, detect_ms, n_detect,
        narrow_ms, n_narrow) -- narrow: the narrow phase of non-spherical bodies."""
        out = np.zeros(9)
        self._check(lib().grains_gpu_profile_steps(self.h, int(n), _d(out)))
        return {"total_ms": out[0], "force_ms": out[1], "n_force": int(out[2]), "bin_ms": out[3],
                "n_bin": int(out[4]), "detect_ms": out[5], "n_detect": int(out[6]), "narrow_ms": out[7],
                "n_narrow": int(out[8])}

    def synchronize(self):
        self._check(lib().grains_gpu_synchronize(self.h))

    def graph_stats(self):
        """(graphs held or -1 if capture is refused, replays, reason)"""
        n, r = ctypes.c_int(), ctypes.c_long()
        why = ctypes.create_string_buffer(256)
        self._check(lib().grains_gpu_graph_stats(self.h, ctypes.byref(n), ctypes.byref(r), why, 256))
        return n.value, r.value, why.value.decode()

    def compute_forces(self):
        self._check(lib().grains_gpu_compute_forces(self.h))

    def force_sums(self):
        """Forces at the current state and, per particle (id order), the sum of
        the magnitudes of its contact forces (grains_gpu_force_sums)."""
        out = np.zeros(self.n)
        self._check(lib().grains_gpu_force_sums(self.h, _d(out)))
        return out

    @property
    def time(self):
        return lib().grains_gpu_time(self.h)

    def set_time(self, t):
        lib().grains_gpu_set_time(self.h, float(t))

    def state(self):
        n = self.n
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "w", "wb", "F", "T")}
        out["q"] = np.zeros((n, 4))
        self._check(lib().grains_gpu_get_state(self.h, *[_d(out[k]) for k in ("x", "v", "w", "wb", "q", "F", "T")]))
        return out

    def set_state(self, x=None, v=None, w=None, q=None, recompute=True):
        keep = [_c64(a) for a in (x, v, w, q)]
        self._check(lib().grains_gpu_set_state(self.h, *[_d(a) for a in keep], 1 if recompute else 0))

    def masses(self):
        n = self.n
        m, I, wgt = np.zeros(n), np.zeros(n), np.zeros((n, 3))
        self._check(lib().grains_gpu_get_masses(self.h, _d(m), _d(I), _d(wgt)))
        return m, I, wgt

    def neighbor_list(self, capacity=32, skin=-1.0):
        """Neighbour candidates per particle (0: the force sweep searches the
        27 cells itself) and skin of the Verlet list (< 0: a tenth of the
        cell size; 0: the list is rebuilt every step).  Before init() only."""
        self._check(lib().grains_gpu_set_neighbor_list(self.h, int(capacity), float(skin)))

    def neighbor_stats(self):
        """(skin of the Verlet list in use or 0, number of list builds)"""
        skin, builds = ctypes.c_double(), ctypes.c_long()
        self._check(lib().grains_gpu_neighbor_stats(self.h, ctypes.byref(skin), ctypes.byref(builds)))
        return skin.value, builds.value

    def contact_stats(self, on=None):
        """With on=True/False: turns the contact statistics on (cleared) or
        off.  Without: dict(overlap_max, overlap_mean, time_overlap_max,
        gjk_iter_mean, n_overlaps, n_evaluations)."""
        if on is not None:
            self._check(lib().grains_gpu_set_contact_stats(self.h, 1 if on else 0))
            return None
        out = np.zeros(6)
        self._check(lib().grains_gpu_contact_stats(self.h, _d(out)))
        return {"overlap_max": out[0], "overlap_mean": out[1], "time_overlap_max": out[2],
                "gjk_iter_mean": out[3], "n_overlaps": out[4], "n_evaluations": out[5]}

    def record_pairs(self, on=True):
        self._check(lib().grains_gpu_set_record_pairs(self.h, 1 if on else 0))

    def pairs(self):
        """Directed contact records (receiver id, partner id) of the last force evaluation."""
        L = lib()
        n = ctypes.c_size_t()
        self._check(L.grains_gpu_get_pairs(self.h, None, 0, ctypes.byref(n)))
        buf = np.zeros(2 * max(n.value, 1), dtype=np.int32)
        self._check(L.grains_gpu_get_pairs(self.h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n.value,
                                           ctypes.byref(n)))
        return buf[: 2 * n.value].reshape(-1, 2).copy()

    def particle_class(self, cls):
        """Class (RawData particle type) of the particles added next."""
        self._check(lib().grains_gpu_set_particle_class(self.h, int(cls)))

    def velocity_max_mean(self):
        """(max |v|, mean |v|) over the particles (AllComponents::ComputeMaxMeanVelocity)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(lib().grains_gpu_velocity_max_mean(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def coordination_numbers(self):
        """Distinct contacting components per particle (id order) at the last
        force evaluation made with pair recording on."""
        out = np.zeros(self.n, dtype=np.int32)
        self._check(lib().grains_gpu_coordination_numbers(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out

    def coordination_at_state(self):
        """Coordination numbers (id order) at the current state, from a force
        evaluation whose torsor is then discarded (the run is unchanged)."""
        out = np.zeros(self.n, dtype=np.int32)
        self._check(lib().grains_gpu_coordination_at_state(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out

    def write_rawdata(self, root, binary=False, truncate=False, coord=None):
        """One RawData frame at the current time (files <root>_*.dat)."""
        cp = None
        if coord is not None:
            coord = np.ascontiguousarray(coord, dtype=np.int32)
            if coord.shape != (self.n,):
                raise ValueError("coord needs one value per particle")
            cp = coord.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        self._check(lib().grains_gpu_write_rawdata(self.h, os.fsencode(root), int(bool(binary)), int(bool(truncate)),
                                                   cp))

    def set_outputs(self, save_times=(), save_root=None, rawdata_root=None, rawdata_binary=False):
        """Output settings of run() (<TimeSave>, <RestartFile Name>, <RawData>)."""
        t = _c64(np.asarray(save_times, dtype=np.float64).reshape(-1))
        self._check(lib().grains_gpu_set_outputs(self.h, _d(t), len(t), os.fsencode(save_root) if save_root else None,
                                                 os.fsencode(rawdata_root) if rawdata_root else None,
                                                 int(bool(rawdata_binary))))

    def construction_state(self):
        """The values the system was built from (id order): dict(x, v, w[, B])."""
        n = self.n = int(lib().grains_gpu_num_particles(self.h))
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "w")}
        B = np.zeros((n, 9))
        if lib().grains_gpu_construction_state(self.h, _d(out["x"]), _d(out["v"]), _d(out["w"]), _d(B)) == 0:
            out["B"] = B.reshape(n, 3, 3)
        else:
            self._check(lib().grains_gpu_construction_state(self.h, _d(out["x"]), _d(out["v"]), _d(out["w"]), None))
        return out

    def try_insert(self):
        """One insertion call of a time step (insertion over time): True if a
        particle went in."""
        k = ctypes.c_int()
        self._check(lib().grains_gpu_try_insert(self.h, ctypes.byref(k)))
        self.n = int(lib().grains_gpu_num_particles(self.h))
        return bool(k.value)

    def pending_insertions(self):
        k = ctypes.c_long()
        self._check(lib().grains_gpu_pending_insertions(self.h, ctypes.byref(k)))
        return k.value

    def run(self, verbose=False):
        """Grains::Simulation to the end time with its outputs."""
        self._check(lib().grains_gpu_run(self.h, int(bool(verbose))))
        self.n = int(lib().grains_gpu_num_particles(self.h))

    def contact_counts(self):
        """(directed contact records, of which on the Coulomb cap) of the last
        force evaluation made with pair recording on."""
        n, ns = ctypes.c_size_t(), ctypes.c_size_t()
        self._check(lib().grains_gpu_contact_counts(self.h, ctypes.byref(n), ctypes.byref(ns)))
        return n.value, ns.value

    def grid(self):
        dims = (ctypes.c_int * 3)()
        edge = np.zeros(3)
        self._check(lib().grains_gpu_grid(self.h, dims, _d(edge)))
        return tuple(dims), edge

    def recoveries(self):
        n = ctypes.c_long()
        self._check(lib().grains_gpu_recoveries(self.h, ctypes.byref(n)))
        return n.value

    def adjoint_forward(self, nsteps, checkpoint_every=50):
        """nsteps steps, keeping the state every checkpoint_every steps for
        adjoint_backward."""
        self._check(lib().grains_gpu_adjoint_forward(self.h, int(nsteps), int(checkpoint_every)))

    def adjoint_backward(self, dJ_dx=None, dJ_dv=None, dJ_dw=None, materials=(0, 0)):
        """Gradients of J with respect to the state at the start of the last
        adjoint_forward run and to (kn, en, muc) of the material pair:
        returns dict x0, v0, w0 ([n, 3]) and params (3,)."""
        n = self.n
        cot = [_c64(a, (n, 3)) for a in (dJ_dx, dJ_dv, dJ_dw)]
        out = {k: np.zeros((n, 3)) for k in ("x0", "v0", "w0")}
        out["params"] = np.zeros(3)
        self._check(lib().grains_gpu_adjoint_backward(self.h, int(materials[0]), int(materials[1]),
                                                      *[_d(a) for a in cot], _d(out["x0"]), _d(out["v0"]),
                                                      _d(out["w0"]), _d(out["params"])))
        return out


# launch shapes of a batch: 0 = flat mode, else the persistent launch's
# workgroup size per member
BATCH_SHAPES = (0, 64, 128, 192, 256, 320, 384, 448, 512)


class GrainsBatch:
    """Independent systems advanced together by one persistent launch.

    ``members`` are GrainsGPU objects that were configured but not
    initialised; their data is copied into the batch."""

    def __init__(self, members, precision=FP64):
        L = lib()
        self.h = ctypes.c_void_p()
        arr = (ctypes.c_void_p * len(members))(*[m.h.value for m in members])
        code = L.grains_gpu_batch_create(arr, len(members), precision, ctypes.byref(self.h))
        if code:
            msg = L.grains_gpu_batch_last_error(self.h).decode() if self.h.value else ""
            raise GrainsError(code, msg)
        self.nsys = len(members)
        self.sizes = [int(L.grains_gpu_batch_num_particles(self.h, i)) for i in range(self.nsys)]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value and _lib is not None:
            _lib.grains_gpu_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        self.close()

    def _check(self, code):
        if code:
            raise GrainsError(code, lib().grains_gpu_batch_last_error(self.h).decode())

    @property
    def n_total(self):
        return int(lib().grains_gpu_batch_num_particles(self.h, -1))

    def set_block(self, block):
        self._check(lib().grains_gpu_batch_set_block(self.h, int(block)))

    def step(self, n, chunk=0):
        self._check(lib().grains_gpu_batch_step(self.h, int(n), int(chunk)))

    def step_timed(self, n, chunk=0):
        ms = ctypes.c_double()
        self._check(lib().grains_gpu_batch_step_timed(self.h, int(n), int(chunk), ctypes.byref(ms)))
        return ms.value

    def autotune(self, trial_steps=20):
        best = ctypes.c_int()
        ms = np.zeros(9)
        self._check(lib().grains_gpu_batch_autotune(self.h, int(trial_steps), ctypes.byref(best), _d(ms)))
        return best.value, dict(zip(BATCH_SHAPES, ms.tolist()))

    def state(self, i):
        n = self.sizes[i]
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "w", "F", "T")}
        out["q"] = np.zeros((n, 4))
        self._check(lib().grains_gpu_batch_get_state(self.h, int(i), *[_d(out[k]) for k in ("x", "v", "w", "q", "F", "T")]))
        return out
