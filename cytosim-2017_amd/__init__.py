"""MI355X implicit mechanics step of Cytosim (a drop-in for `Meca`).

Python binding of the C ABI in include/cymeca.h (libcymeca.so, HIP for
gfx950).  The product path is the HIP library: importing this module never
falls back to any CPU implementation; a missing library raises.

The directory name is not a Python identifier; load it with `load()` from
`__graft_entry__` or with importlib.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "_build", "libcymeca.so")

OK, ERR_INVALID, ERR_HIP, ERR_PROJECTION, ERR_CONVERGENCE, ERR_UNSUPPORTED = range(6)
MAX_FIBER_POINTS = 512
FRAME_FLOAT, FRAME_DOUBLE = 4, 8
TOL_REFERENCE, TOL_RELATIVE = 0, 1
PRECOND_NONE, PRECOND_FP64, PRECOND_FP32 = 0, 1, 2
GUESS_ZERO, GUESS_PREVIOUS, GUESS_LEARNED = 0, 1, 2
VEC_PTS, VEC_SOL, VEC_BAS, VEC_RHS, VEC_FOR, VEC_PREV = range(6)

EXPORTED = [
    "cymg_create", "cymg_destroy", "cymg_last_error", "cymg_set_params", "cymg_clear",
    "cymg_add_fiber", "cymg_add_fiber_ex", "cymg_set_identities", "cymg_add_solid", "cymg_get_solid_timers", "cymg_nb_points", "cymg_set_systems", "cymg_get_system_reports",
    "cymg_prepare", "cymg_add_pure_force",
    "cymg_inter_link_ee", "cymg_inter_link_ie", "cymg_inter_link_ii", "cymg_inter_long_link_ee",
    "cymg_inter_clamp_e", "cymg_inter_clamp_i", "cymg_inter_long_clamp", "cymg_inter_plane_e",
    "cymg_add_pure_force_i", "cymg_inter_plane_i", "cymg_inter_long_link_ie", "cymg_inter_long_link_ii",
    "cymg_inter_side_link_ie", "cymg_inter_side_link_ii", "cymg_inter_side_side_link", "cymg_inter_sliding_link_ie",
    "cymg_inter_sliding_link_ii", "cymg_inter_side_sliding_link_ie", "cymg_inter_side_sliding_link_ii",
    "cymg_inter_torque", "cymg_inter_side_clamp", "cymg_inter_sliding_clamp",
    "cymg_set_steric", "cymg_steric_interactions", "cymg_set_frame_records", "cymg_write_frame", "cymg_read_frame",
    "cymg_confine_sphere", "cymg_set_couples", "cymg_couples_interactions", "cymg_set_couple_state",
    "cymg_get_couples", "cymg_step_couples", "cymg_nearby_segments", "cymg_nearby_segments_noalloc", "cymg_set_aster_links",
    "cymg_aster_interactions", "cymg_solve",
    "cymg_setup", "cymg_iterate", "cymg_apply", "cymg_multiply", "cymg_compute_preconditioner",
    "cymg_precondition", "cymg_compute_forces", "cymg_multiply_device", "cymg_precondition_device",
    "cymg_get_vector", "cymg_dump", "cymg_get_points", "cymg_get_tensions", "cymg_get_block", "cymg_get_matrix",
    "cymg_get_timings", "cymg_enable_timing", "cymg_time_kernel", "cymg_set_option", "cymg_get_option", "cymg_autotune",
    "cymg_guess_init", "cymg_guess_weights", "cymg_guess_record", "cymg_guess_samples",
    "cymg_guess_train", "cymg_guess_eval", "cymg_guess_check", "cymg_guess_features", "cymg_guess_predict",
]
OPT_KRYLOV_LAUNCHES, OPT_BLOCK_INVERSE, OPT_TILE_POINTS, OPT_GRAPHS, OPT_TILE_THREADS, OPT_GUESS_GUARD, \
    OPT_FP32_APPLY, OPT_BLOCK_STAGE = range(1, 9)
GUESS_INPUTS, GUESS_OUTPUTS, GUESS_POINTS, GUESS_WEIGHTS = 384, 96, 31, 254816


class SimParams(C.Structure):
    _fields_ = [("time_step", C.c_double), ("kT", C.c_double), ("viscosity", C.c_double),
                ("hydro_radius0", C.c_double), ("hydro_radius1", C.c_double)]


class FiberDesc(C.Structure):
    _fields_ = [("nb_points", C.c_uint32), ("segmentation", C.c_double), ("rigidity", C.c_double),
                ("hydro_radius0", C.c_double), ("hydro_radius1", C.c_double), ("surface_effect", C.c_int),
                ("cylinder_height", C.c_double)]


class SolveParams(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("tol_mode", C.c_int), ("precondition", C.c_int),
                ("max_iter", C.c_int), ("guess", C.c_int), ("fallback", C.c_int),
                ("gauss_on_device", C.c_int)]


class SolveReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("flag", C.c_int), ("converged", C.c_int),
                ("fallback", C.c_int), ("precond_used", C.c_int), ("total_iterations", C.c_int),
                ("residual", C.c_double), ("threshold", C.c_double), ("noise_level", C.c_double),
                ("guess_used", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Timings(C.Structure):
    _fields_ = [("assemble_ms", C.c_double), ("setup_ms", C.c_double), ("precond_ms", C.c_double),
                ("iterate_ms", C.c_double), ("apply_ms", C.c_double), ("matvec_avg_ms", C.c_double),
                ("pcapply_avg_ms", C.c_double), ("matvec_count", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MotorProps(C.Structure):
    _fields_ = [("max_speed", C.c_double), ("stall_force", C.c_double), ("limit_speed", C.c_int),
                ("hold_growing_end", C.c_int)]


class CymgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ConvergenceFailure(CymgError):
    """Exception("convergence failure in solver") of the reference (meca.cc:795)."""


class ProjectionFailure(CymgError):
    """Exception("could not build Fiber projection") (rigid_fiber_project.cc:84-85)."""


def build(verbose=False):
    """Compile the HIP library for gfx950 in-tree (hipcc via the package Makefile)."""
    jobs = os.environ.get("MAX_JOBS", "8")
    out = subprocess.run(["make", "-j", jobs, "-C", HERE], capture_output=not verbose, text=True)
    if out.returncode:
        raise RuntimeError("HIP build failed:\n" + (out.stdout or "") + (out.stderr or ""))


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libcymeca.so not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    u32 = C.c_uint32
    sig = {
        "cymg_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(vp)]),
        "cymg_destroy": (C.c_int, [vp]),
        "cymg_last_error": (C.c_char_p, [vp]),
        "cymg_set_params": (C.c_int, [vp, C.POINTER(SimParams)]),
        "cymg_clear": (C.c_int, [vp]),
        "cymg_add_fiber": (C.c_int, [vp, u32, C.c_double, C.c_double, C.POINTER(u32)]),
        "cymg_add_fiber_ex": (C.c_int, [vp, C.POINTER(FiberDesc), C.POINTER(u32)]),
        "cymg_set_identities": (C.c_int, [vp, u32, vp]),
        "cymg_add_solid": (C.c_int, [vp, u32, vp, vp, C.c_int, C.POINTER(u32)]),
        "cymg_get_solid_timers": (C.c_int, [vp, u32, vp]),
        "cymg_nb_points": (C.c_int, [vp, C.POINTER(u32)]),
        "cymg_set_systems": (C.c_int, [vp, u32, vp]),
        "cymg_get_system_reports": (C.c_int, [vp, u32, vp]),
        "cymg_prepare": (C.c_int, [vp, vp]),
        "cymg_add_pure_force": (C.c_int, [vp, u32, dp]),
        "cymg_inter_link_ee": (C.c_int, [vp, u32, u32, C.c_double]),
        "cymg_inter_link_ie": (C.c_int, [vp, u32, u32, C.c_double, u32, C.c_double]),
        "cymg_inter_link_ii": (C.c_int, [vp, u32, u32, C.c_double, u32, u32, C.c_double, C.c_double]),
        "cymg_inter_long_link_ee": (C.c_int, [vp, u32, u32, C.c_double, C.c_double]),
        "cymg_inter_clamp_e": (C.c_int, [vp, u32, dp, C.c_double]),
        "cymg_inter_clamp_i": (C.c_int, [vp, u32, u32, C.c_double, dp, C.c_double]),
        "cymg_inter_long_clamp": (C.c_int, [vp, u32, dp, C.c_double, C.c_double]),
        "cymg_inter_plane_e": (C.c_int, [vp, u32, dp, dp, C.c_double]),
        "cymg_add_pure_force_i": (C.c_int, [vp, u32, u32, C.c_double, dp]),
        "cymg_inter_plane_i": (C.c_int, [vp, u32, u32, C.c_double, dp, dp, C.c_double]),
        "cymg_dump": (C.c_int, [vp, C.c_char_p]),
        "cymg_set_steric": (C.c_int, [vp, C.POINTER(StericParams), u32, vp, vp, vp]),
        "cymg_set_frame_records": (C.c_int, [vp, u32, vp]),
        "cymg_write_frame": (C.c_int, [vp, C.POINTER(FrameParams), vp, C.c_size_t, C.POINTER(C.c_size_t)]),
        "cymg_read_frame": (C.c_int, [vp, C.c_size_t, C.c_int, vp, u32, C.POINTER(C.c_uint32), vp, vp, C.c_size_t,
                                      C.POINTER(C.c_size_t), C.POINTER(C.c_double), C.POINTER(C.c_size_t)]),
        "cymg_steric_interactions": (C.c_int, [vp]),
        "cymg_inter_long_link_ie": (C.c_int, [vp, u32, u32, C.c_double, u32, C.c_double, C.c_double]),
        "cymg_inter_long_link_ii": (C.c_int, [vp, u32, u32, C.c_double, u32, u32, C.c_double, C.c_double, C.c_double]),
        "cymg_inter_side_link_ie": (C.c_int, [vp, u32, u32, C.c_double, u32, C.c_double, C.c_double]),
        "cymg_inter_side_link_ii": (C.c_int, [vp, u32, u32, C.c_double, u32, u32, C.c_double, C.c_double, C.c_double]),
        "cymg_inter_side_side_link": (C.c_int, [vp, u32, u32, C.c_double, u32, u32, C.c_double, C.c_double, C.c_double]),
        "cymg_inter_sliding_link_ie": (C.c_int, [vp, u32, u32, C.c_double, u32, C.c_double]),
        "cymg_inter_sliding_link_ii": (C.c_int, [vp, u32, u32, C.c_double, u32, u32, C.c_double, C.c_double]),
        "cymg_inter_side_sliding_link_ie": (C.c_int, [vp, u32, u32, C.c_double, u32, C.c_double, C.c_double]),
        "cymg_inter_side_sliding_link_ii": (C.c_int, [vp, u32, u32, C.c_double, u32, u32, C.c_double, C.c_double, C.c_double]),
        "cymg_inter_torque": (C.c_int, [vp, u32, u32, C.c_double, u32, u32, C.c_double, C.c_double, C.c_double, C.c_double]),
        "cymg_inter_side_clamp": (C.c_int, [vp, u32, u32, C.c_double, dp, C.c_double, C.c_double]),
        "cymg_inter_sliding_clamp": (C.c_int, [vp, u32, u32, C.c_double, dp, C.c_double]),
        "cymg_confine_sphere": (C.c_int, [vp, C.c_double, C.c_double]),
        "cymg_set_couples": (C.c_int, [vp, u32, vp, vp, vp, vp, vp]),
        "cymg_couples_interactions": (C.c_int, [vp]),
        "cymg_set_couple_state": (C.c_int, [vp, vp]),
        "cymg_get_couples": (C.c_int, [vp, vp, vp, vp]),
        "cymg_step_couples": (C.c_int, [vp, C.POINTER(MotorProps)]),
        "cymg_nearby_segments": (C.c_int, [vp, u32, vp, vp, C.c_double, C.c_int, vp, u32, vp, vp, vp, vp,
                                           C.POINTER(u32)]),
        "cymg_nearby_segments_noalloc": (C.c_int, [vp, u32, vp, vp, C.c_double, C.c_int, vp, u32, vp, vp, vp, vp,
                                                   C.POINTER(u32)]),
        "cymg_set_aster_links": (C.c_int, [vp, u32, vp, vp, vp, vp, vp, vp]),
        "cymg_aster_interactions": (C.c_int, [vp]),
        "cymg_solve": (C.c_int, [vp, C.POINTER(SolveParams), vp, C.POINTER(SolveReport)]),
        "cymg_setup": (C.c_int, [vp, vp, C.c_int]),
        "cymg_iterate": (C.c_int, [vp, C.POINTER(SolveParams), C.POINTER(SolveReport)]),
        "cymg_apply": (C.c_int, [vp]),
        "cymg_multiply": (C.c_int, [vp, vp, vp]),
        "cymg_compute_preconditioner": (C.c_int, [vp, C.c_int]),
        "cymg_precondition": (C.c_int, [vp, vp, vp]),
        "cymg_compute_forces": (C.c_int, [vp]),
        "cymg_multiply_device": (C.c_int, [vp, vp, vp]),
        "cymg_precondition_device": (C.c_int, [vp, vp, vp]),
        "cymg_get_vector": (C.c_int, [vp, C.c_int, vp]),
        "cymg_get_points": (C.c_int, [vp, vp]),
        "cymg_get_tensions": (C.c_int, [vp, vp]),
        "cymg_get_block": (C.c_int, [vp, u32, C.c_int, vp]),
        "cymg_get_matrix": (C.c_int, [vp, C.c_int, C.POINTER(u32), vp, vp, vp]),
        "cymg_get_timings": (C.c_int, [vp, C.POINTER(Timings)]),
        "cymg_enable_timing": (C.c_int, [vp, C.c_int]),
        "cymg_time_kernel": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "cymg_set_option": (C.c_int, [vp, C.c_int, C.c_int]),
        "cymg_get_option": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int)]),
        "cymg_autotune": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "cymg_guess_init": (C.c_int, [vp, C.c_uint64]),
        "cymg_guess_weights": (C.c_int, [vp, C.c_int, vp]),
        "cymg_guess_record": (C.c_int, [vp, u32]),
        "cymg_guess_samples": (C.c_int, [vp, C.POINTER(u32)]),
        "cymg_guess_train": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, C.c_uint64, C.POINTER(C.c_double)]),
        "cymg_guess_eval": (C.c_int, [vp, C.c_int, C.c_uint64, C.POINTER(C.c_double)]),
        "cymg_guess_check": (C.c_int, [vp, vp, vp, C.c_int, C.POINTER(C.c_double), vp]),
        "cymg_guess_features": (C.c_int, [vp, vp, vp]),
        "cymg_guess_predict": (C.c_int, [vp, C.c_int, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class StericParams(C.Structure):
    _fields_ = [("extension", C.c_double * 3), ("max_range", C.c_double), ("stiff_push", C.c_double),
                ("stiff_pull", C.c_double)]


class FrameRecord(C.Structure):
    _fields_ = [("number", C.c_uint32), ("mark", C.c_int32), ("prop", C.c_uint32), ("signature", C.c_uint32),
                ("cut_wished", C.c_double), ("abscissa", C.c_double)]


class FrameParams(C.Structure):
    _fields_ = [("time", C.c_double), ("date", C.c_char_p), ("real_bytes", C.c_int), ("before", C.c_void_p),
                ("nbefore", C.c_size_t), ("after", C.c_void_p), ("nafter", C.c_size_t)]


class FrameObject(C.Structure):
    _fields_ = [("tag", C.c_char), ("prop", C.c_uint32), ("number", C.c_uint32), ("mark", C.c_int32),
                ("signature", C.c_uint32), ("length", C.c_double), ("cut_wished", C.c_double),
                ("abscissa", C.c_double), ("points", C.c_uint32), ("first", C.c_uint64)]


def read_frame(data, dim, max_objects=1 << 16, max_points=1 << 22):
    """cymg_read_frame: the first frame of `data` (bytes) -> dict with time,
    used (bytes up to the frame's end), objects (list of dicts) and the points
    [npoints, dim] and radii of the fiber and solid sections"""
    L = lib()
    buf = C.create_string_buffer(bytes(data), len(data))
    objs = (FrameObject * max_objects)()
    pts = np.zeros((max_points, dim))
    rad = np.zeros(max_points)
    no, npt, used, t = C.c_uint32(), C.c_size_t(), C.c_size_t(), C.c_double()
    rc = L.cymg_read_frame(buf, len(data), dim, objs, max_objects, C.byref(no), _p(pts), _p(rad), max_points,
                           C.byref(npt), C.byref(t), C.byref(used))
    if rc:
        raise CymgError(rc, L.cymg_last_error(None).decode())
    if no.value > max_objects or npt.value > max_points:
        raise CymgError(ERR_INVALID, "frame larger than the buffers")
    out = []
    for o in objs[:no.value]:
        out.append(dict(tag=o.tag.decode(), prop=o.prop, number=o.number, mark=o.mark, signature=o.signature,
                        length=o.length, cut_wished=o.cut_wished, abscissa=o.abscissa, points=o.points,
                        first=o.first))
    return dict(time=t.value, used=used.value, objects=out, points=pts[:npt.value].copy(),
                radii=rad[:npt.value].copy())


def _vec3(v, dim):
    a = np.zeros(3)
    a[:dim] = np.asarray(v, dtype=np.float64)[:dim]
    return a.ctypes.data_as(C.POINTER(C.c_double)), a


class Meca:
    """One GPU Meca instance: the reference's `Meca` interface (meca.h:160-327)
    over the C ABI.  Errors raise CymgError; the solver's failure raises
    ConvergenceFailure as the reference throws."""

    def __init__(self, dim, device=0, params=None):
        self.L = lib()
        self.dim = dim
        h = C.c_void_p()
        rc = self.L.cymg_create(dim, device, C.byref(h))
        if rc:
            raise CymgError(rc, "cymg_create failed (dim must be 2 or 3, and a GPU must be visible)")
        self.h = h
        self.np = []
        self._ncouples = 0
        if params is not None:
            self.set_params(**params)

    def close(self):
        if getattr(self, "h", None):
            self.L.cymg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc:
            msg = self.L.cymg_last_error(self.h).decode()
            if rc == ERR_CONVERGENCE:
                raise ConvergenceFailure(rc, msg)
            if rc == ERR_PROJECTION:
                raise ProjectionFailure(rc, msg)
            raise CymgError(rc, msg)

    # registration ------------------------------------------------------------
    def set_params(self, time_step, kT, viscosity, hydro_radius0=0.0125, hydro_radius1=10.0):
        p = SimParams(time_step, kT, viscosity, hydro_radius0, hydro_radius1)
        self._chk(self.L.cymg_set_params(self.h, C.byref(p)))

    def clear(self):
        self._chk(self.L.cymg_clear(self.h))
        self.np = []

    def add_fiber(self, nb_points, segmentation, rigidity, hydro_radius0=None, hydro_radius1=None,
                  surface_height=None):
        """Meca::add of a fiber; optional per-fiber FiberProp drag settings
        (hydrodynamic radii, surface_height > 0 for the surface drag model)."""
        mi = C.c_uint32()
        if hydro_radius0 is None and hydro_radius1 is None and surface_height is None:
            self._chk(self.L.cymg_add_fiber(self.h, int(nb_points), float(segmentation), float(rigidity),
                                            C.byref(mi)))
        else:
            d = FiberDesc(int(nb_points), float(segmentation), float(rigidity), float(hydro_radius0 or 0.0),
                          float(hydro_radius1 or 0.0), int(bool(surface_height)), float(surface_height or 0.0))
            self._chk(self.L.cymg_add_fiber_ex(self.h, C.byref(d), C.byref(mi)))
        self.np.append(int(nb_points))
        return mi.value

    def set_identities(self, ids):
        """Inventoried::number() of every Mecable (fibers, then Solids): the
        device then carries each object's points and previous solution over a
        new registration order."""
        a = np.ascontiguousarray(ids, dtype=np.uint32)
        self._chk(self.L.cymg_set_identities(self.h, a.size, _p(a)))

    def add_solid(self, radius, points, reshape_timer=0):
        """A Solid (after every fiber): per-point radii and its reference shape."""
        r = np.ascontiguousarray(radius, dtype=np.float64)
        pts = np.ascontiguousarray(points, dtype=np.float64).ravel()
        assert pts.size == self.dim * r.size
        mi = C.c_uint32()
        self._chk(self.L.cymg_add_solid(self.h, r.size, _p(r), _p(pts), int(reshape_timer), C.byref(mi)))
        self.np.append(int(r.size))
        self.nsol = getattr(self, "nsol", 0) + 1
        return mi.value

    def solid_timers(self):
        n = getattr(self, "nsol", 0)
        out = np.zeros(n, np.int32)
        self._chk(self.L.cymg_get_solid_timers(self.h, n, _p(out) if n else None))
        return out

    def set_systems(self, first_fiber):
        ff = np.ascontiguousarray(first_fiber, dtype=np.uint32)
        self._chk(self.L.cymg_set_systems(self.h, ff.size, _p(ff)))
        self.nsys = int(ff.size)

    def system_reports(self):
        n = getattr(self, "nsys", 1)
        arr = (SolveReport * n)()
        self._chk(self.L.cymg_get_system_reports(self.h, n, C.cast(arr, C.c_void_p)))
        return [arr[i].as_dict() for i in range(n)]

    @property
    def nb_points(self):
        return int(sum(self.np))

    @property
    def size(self):
        return self.dim * self.nb_points

    def prepare(self, points=None):
        if points is None:
            self._chk(self.L.cymg_prepare(self.h, None))
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).ravel()
            assert pts.size == self.size
            self._chk(self.L.cymg_prepare(self.h, _p(pts)))

    # assembly ------------------------------------------------------------------
    def inter_link_ii(self, a1, a2, ca, b1, b2, cb, w):
        self._chk(self.L.cymg_inter_link_ii(self.h, a1, a2, ca, b1, b2, cb, w))

    def inter_link_ie(self, a1, a2, ca, b, w):
        self._chk(self.L.cymg_inter_link_ie(self.h, a1, a2, ca, b, w))

    def inter_link_ee(self, a, b, w):
        self._chk(self.L.cymg_inter_link_ee(self.h, a, b, w))

    def inter_long_link_ee(self, a, b, length, w):
        self._chk(self.L.cymg_inter_long_link_ee(self.h, a, b, length, w))

    def inter_clamp_e(self, a, g, w):
        gp, keep = _vec3(g, self.dim)
        self._chk(self.L.cymg_inter_clamp_e(self.h, a, gp, w))

    def inter_clamp_i(self, a1, a2, ca, g, w):
        gp, keep = _vec3(g, self.dim)
        self._chk(self.L.cymg_inter_clamp_i(self.h, a1, a2, ca, gp, w))

    def inter_long_clamp(self, a, center, length, w):
        cp, keep = _vec3(center, self.dim)
        self._chk(self.L.cymg_inter_long_clamp(self.h, a, cp, length, w))

    def inter_plane_e(self, a, direction, g, w):
        dp, k1 = _vec3(direction, self.dim)
        gp, k2 = _vec3(g, self.dim)
        self._chk(self.L.cymg_inter_plane_e(self.h, a, dp, gp, w))

    def add_pure_force(self, a, force):
        fp, keep = _vec3(force, self.dim)
        self._chk(self.L.cymg_add_pure_force(self.h, a, fp))

    def add_pure_force_i(self, a1, a2, ca, force):
        """Meca::addPureForce(PointInterpolated) (meca_inter.cc:140-156)"""
        fp, keep = _vec3(force, self.dim)
        self._chk(self.L.cymg_add_pure_force_i(self.h, a1, a2, ca, fp))

    def inter_plane_i(self, a1, a2, ca, direction, g, w):
        """Meca::interPlane(PointInterpolated) (meca_inter.cc:2340-2378)"""
        dp, k1 = _vec3(direction, self.dim)
        gp, k2 = _vec3(g, self.dim)
        self._chk(self.L.cymg_inter_plane_i(self.h, a1, a2, ca, dp, gp, w))

    # Interactions that read the positions: evaluated on the device from the
    # points of prepare() when the matrices are built.

    def inter_long_link_ie(self, a1, a2, ca, b, length, w):
        """Meca::interLongLink(PointInterpolated, PointExact) (meca_inter.cc:425-505)"""
        self._chk(self.L.cymg_inter_long_link_ie(self.h, a1, a2, ca, b, length, w))

    def inter_long_link_ii(self, a1, a2, ca, b1, b2, cb, length, w):
        """Meca::interLongLink(PointInterpolated, PointInterpolated) (meca_inter.cc:516-596)"""
        self._chk(self.L.cymg_inter_long_link_ii(self.h, a1, a2, ca, b1, b2, cb, length, w))

    def inter_side_link_ie(self, a1, a2, ca, b, length, w):
        """Meca::interSideLink(PointInterpolated, PointExact) (meca_inter.cc:617-759)"""
        self._chk(self.L.cymg_inter_side_link_ie(self.h, a1, a2, ca, b, length, w))

    def inter_side_link_ii(self, a1, a2, ca, b1, b2, cb, length, w):
        """Meca::interSideLink(PointInterpolated, PointInterpolated) (meca_inter.cc:764-935)"""
        self._chk(self.L.cymg_inter_side_link_ii(self.h, a1, a2, ca, b1, b2, cb, length, w))

    def inter_side_side_link(self, a1, a2, ca, b1, b2, cb, length, w):
        """Meca::interSideSideLink (meca_inter.cc:943-1059), 2D only"""
        self._chk(self.L.cymg_inter_side_side_link(self.h, a1, a2, ca, b1, b2, cb, length, w))

    def inter_sliding_link_ie(self, a1, a2, ca, b, w):
        """Meca::interSlidingLink(PointInterpolated, PointExact) (meca_inter.cc:1075-1140)"""
        self._chk(self.L.cymg_inter_sliding_link_ie(self.h, a1, a2, ca, b, w))

    def inter_sliding_link_ii(self, a1, a2, ca, b1, b2, cb, w):
        """Meca::interSlidingLink(PointInterpolated, PointInterpolated) (meca_inter.cc:1153-1258)"""
        self._chk(self.L.cymg_inter_sliding_link_ii(self.h, a1, a2, ca, b1, b2, cb, w))

    def inter_side_sliding_link_ie(self, a1, a2, ca, b, length, w):
        """Meca::interSideSlidingLink(PointInterpolated, PointExact) (meca_inter.cc:1266-1499)"""
        self._chk(self.L.cymg_inter_side_sliding_link_ie(self.h, a1, a2, ca, b, length, w))

    def inter_side_sliding_link_ii(self, a1, a2, ca, b1, b2, cb, length, w):
        """Meca::interSideSlidingLink(PointInterpolated, PointInterpolated) (meca_inter.cc:1508-1751)"""
        self._chk(self.L.cymg_inter_side_sliding_link_ii(self.h, a1, a2, ca, b1, b2, cb, length, w))

    def inter_torque(self, a1, a2, ca, b1, b2, cb, cosinus, sinus, w):
        """Meca::interTorque2D (meca_inter.cc:1772-1869), 2D only"""
        self._chk(self.L.cymg_inter_torque(self.h, a1, a2, ca, b1, b2, cb, cosinus, sinus, w))

    def inter_side_clamp(self, a1, a2, ca, g, length, w):
        """Meca::interSideClamp (meca_inter.cc:2032-2226)"""
        gp, keep = _vec3(g, self.dim)
        self._chk(self.L.cymg_inter_side_clamp(self.h, a1, a2, ca, gp, length, w))

    def inter_sliding_clamp(self, a1, a2, ca, g, w):
        """Meca::interSlidingClamp (meca_inter.cc:2241-2283)"""
        gp, keep = _vec3(g, self.dim)
        self._chk(self.L.cymg_inter_sliding_clamp(self.h, a1, a2, ca, gp, w))

    def set_steric(self, extension, max_range, push, pull, steric, radius, range_):
        """steric parameters (SimulProp::steric_max_range, steric_stiffness_push/pull,
        Space::extension) and, per Mecable in registration order, the steric flag
        and the fibers' steric_radius / steric_range"""
        sp = StericParams()
        for d in range(3):
            sp.extension[d] = extension[d] if d < len(extension) else 0.0
        sp.max_range, sp.stiff_push, sp.stiff_pull = max_range, push, pull
        st = np.ascontiguousarray(steric, dtype=np.int32)
        rd = np.ascontiguousarray(radius, dtype=np.float64)
        rg = np.ascontiguousarray(range_, dtype=np.float64)
        self._chk(self.L.cymg_set_steric(self.h, C.byref(sp), st.size, _p(st), _p(rd), _p(rg)))

    def set_frame_records(self, records):
        """per Mecable in registration order: (number, mark, prop, signature,
        cut_wished, abscissa) of its trajectory record"""
        arr = (FrameRecord * max(1, len(records)))()
        for k, r in enumerate(records):
            arr[k] = FrameRecord(*r)
        self._chk(self.L.cymg_set_frame_records(self.h, len(records), arr))

    def write_frame(self, time, date, real_bytes=FRAME_FLOAT, before=b"", after=b""):
        """Simul::writeObjects (simul_file.cc:402-489) in binary mode, the fiber and
        solid sections packed on the device; returns the frame's bytes"""
        bb = C.create_string_buffer(bytes(before), max(1, len(before)))
        ab = C.create_string_buffer(bytes(after), max(1, len(after)))
        fp = FrameParams(time, date.encode() if isinstance(date, str) else date, real_bytes,
                         C.cast(bb, C.c_void_p), len(before), C.cast(ab, C.c_void_p), len(after))
        n = C.c_size_t()
        self._chk(self.L.cymg_write_frame(self.h, C.byref(fp), None, 0, C.byref(n)))
        out = C.create_string_buffer(n.value)
        self._chk(self.L.cymg_write_frame(self.h, C.byref(fp), out, n.value, C.byref(n)))
        return out.raw[:n.value]

    def steric_interactions(self):
        """Simul::setStericInteractions (simul_solve.cc:139-188) on the device"""
        self._chk(self.L.cymg_steric_interactions(self.h))

    def confine_sphere(self, radius, stiffness):
        self._chk(self.L.cymg_confine_sphere(self.h, radius, stiffness))

    def set_couples(self, fiber_a, abs_a, fiber_b, abs_b, stiffness):
        fa = np.ascontiguousarray(fiber_a, dtype=np.uint32)
        fb = np.ascontiguousarray(fiber_b, dtype=np.uint32)
        aa = np.ascontiguousarray(abs_a, dtype=np.float64)
        ab = np.ascontiguousarray(abs_b, dtype=np.float64)
        w = np.ascontiguousarray(stiffness, dtype=np.float64)
        self._chk(self.L.cymg_set_couples(self.h, fa.size, _p(fa), _p(aa), _p(fb), _p(ab), _p(w)))
        self._ncouples = int(fa.size)

    def couples_interactions(self):
        self._chk(self.L.cymg_couples_interactions(self.h))

    def set_couple_state(self, state):
        """Per uploaded couple: 1 hand 1 attached, 2 hand 2 attached, 4 hand 1
        is a motor, 8 hand 2 is a motor."""
        st = np.ascontiguousarray(state, dtype=np.uint8)
        self._chk(self.L.cymg_set_couple_state(self.h, _p(st)))

    def get_couples(self):
        """(abscissa of hand 1, abscissa of hand 2, state) of the uploaded couples."""
        n = self._ncouples
        aa, ab, st = np.empty(n), np.empty(n), np.empty(n, np.uint8)
        if n:
            self._chk(self.L.cymg_get_couples(self.h, _p(aa), _p(ab), _p(st)))
        return aa, ab, st

    def step_couples(self, max_speed, stall_force, limit_speed=True, hold_growing_end=False):
        """Couple::stepAA/AF/FA with the motor hands, from the current points."""
        mp = MotorProps(float(max_speed), float(stall_force), int(bool(limit_speed)), int(bool(hold_growing_end)))
        self._chk(self.L.cymg_step_couples(self.h, C.byref(mp)))

    def nearby_segments(self, places, range_, inclusive=False, exclude=None, noalloc=False):
        """FiberGrid::nearbySegments for many places (fiber_grid.cc:302-337):
        counts per place and the concatenated (fiber, point, abscissa,
        dist_sqr) of the kept segments, place after place, each in fiber
        order."""
        w = np.ascontiguousarray(places, dtype=np.float64).reshape(-1, self.dim)
        nq = w.shape[0]
        ex = None if exclude is None else np.ascontiguousarray(exclude, dtype=np.uint32)
        counts = np.zeros(max(nq, 1), np.uint32)
        total = C.c_uint32(0)
        cap = 8 * nq + 64          # usually enough: one call; otherwise a second with the exact size
        out = None
        for _ in range(2):
            f, p = np.zeros(max(cap, 1), np.uint32), np.zeros(max(cap, 1), np.uint32)
            a, d = np.zeros(max(cap, 1)), np.zeros(max(cap, 1))
            fn = self.L.cymg_nearby_segments_noalloc if noalloc else self.L.cymg_nearby_segments
            self._chk(fn(self.h, nq, _p(w), _p(ex), float(range_), int(bool(inclusive)), _p(counts), cap, _p(f),
                         _p(p), _p(a), _p(d), C.byref(total)))
            if total.value <= cap:
                out = (f[:total.value], p[:total.value], a[:total.value], d[:total.value])
                break
            cap = total.value
        return counts[:nq], out

    def set_aster_links(self, kind, s1, s2, fiber, abscissa, stiffness):
        kd = np.ascontiguousarray(kind, dtype=np.int32)
        a = np.ascontiguousarray(s1, dtype=np.uint32)
        b = np.ascontiguousarray(s2, dtype=np.uint32)
        f = np.ascontiguousarray(fiber, dtype=np.uint32)
        ab = np.ascontiguousarray(abscissa, dtype=np.float64)
        w = np.ascontiguousarray(stiffness, dtype=np.float64)
        self._chk(self.L.cymg_set_aster_links(self.h, kd.size, _p(kd), _p(a), _p(b), _p(f), _p(ab), _p(w)))

    def aster_interactions(self):
        self._chk(self.L.cymg_aster_interactions(self.h))

    # solve ------------------------------------------------------------------------
    @staticmethod
    def _params(tolerance, tol_mode, precondition, max_iter, guess, fallback, gauss_on_device=0):
        return SolveParams(tolerance, tol_mode, precondition, max_iter, guess, fallback, gauss_on_device)

    def setup(self, gauss=None, gauss_device_ptr=None):
        if gauss_device_ptr is not None:
            self._chk(self.L.cymg_setup(self.h, C.c_void_p(gauss_device_ptr), 1))
        else:
            g = None if gauss is None else np.ascontiguousarray(gauss, dtype=np.float64)
            self._chk(self.L.cymg_setup(self.h, _p(g), 0))

    def iterate(self, tolerance, tol_mode=TOL_REFERENCE, precondition=PRECOND_FP64, max_iter=0,
                guess=GUESS_ZERO, fallback=1):
        sp = self._params(tolerance, tol_mode, precondition, max_iter, guess, fallback)
        rep = SolveReport()
        self._chk(self.L.cymg_iterate(self.h, C.byref(sp), C.byref(rep)))
        return rep.as_dict()

    def apply(self):
        self._chk(self.L.cymg_apply(self.h))

    def solve(self, tolerance, gauss=None, tol_mode=TOL_REFERENCE, precondition=PRECOND_FP64,
              max_iter=0, guess=GUESS_ZERO, fallback=1, gauss_device_ptr=None):
        on_dev = 1 if gauss_device_ptr is not None else 0
        sp = self._params(tolerance, tol_mode, precondition, max_iter, guess, fallback, on_dev)
        rep = SolveReport()
        if on_dev:
            g = C.c_void_p(gauss_device_ptr)
        else:
            garr = None if gauss is None else np.ascontiguousarray(gauss, dtype=np.float64)
            g = _p(garr)
        self._chk(self.L.cymg_solve(self.h, C.byref(sp), g, C.byref(rep)))
        return rep.as_dict()

    # operators ----------------------------------------------------------------------
    def multiply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self._chk(self.L.cymg_multiply(self.h, _p(x), _p(y)))
        return y

    def multiply_device(self, x_ptr, y_ptr):
        self._chk(self.L.cymg_multiply_device(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr)))

    def precondition_device(self, x_ptr, y_ptr):
        self._chk(self.L.cymg_precondition_device(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr)))

    def compute_preconditioner(self, precision=PRECOND_FP64):
        self._chk(self.L.cymg_compute_preconditioner(self.h, precision))

    def precondition(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self._chk(self.L.cymg_precondition(self.h, _p(x), _p(y)))
        return y

    def compute_forces(self):
        self._chk(self.L.cymg_compute_forces(self.h))

    # export --------------------------------------------------------------------------
    def vector(self, which):
        out = np.empty(self.size)
        self._chk(self.L.cymg_get_vector(self.h, which, _p(out)))
        return out

    def fiber_points(self):
        out = np.empty(self.size)
        self._chk(self.L.cymg_get_points(self.h, _p(out)))
        return out

    def export_points(self, out):
        """getPoints of every Mecable into a preallocated float64 array."""
        assert out.dtype == np.float64 and out.size == self.size and out.flags.c_contiguous
        self._chk(self.L.cymg_get_points(self.h, _p(out)))

    def export_vector(self, which, out):
        assert out.dtype == np.float64 and out.size == self.size and out.flags.c_contiguous
        self._chk(self.L.cymg_get_vector(self.h, which, _p(out)))

    def tensions(self):
        nfib = len(self.np) - getattr(self, "nsol", 0)
        out = np.empty(sum(self.np[:nfib]) - nfib)
        self._chk(self.L.cymg_get_tensions(self.h, _p(out)))
        return out

    def get_block(self, f, inverse=False):
        bs = self.dim * self.np[f]
        out = np.empty(bs * bs)
        self._chk(self.L.cymg_get_block(self.h, f, int(bool(inverse)), _p(out)))
        return out.reshape(bs, bs).T.copy()

    def dump(self, directory=None):
        """Meca::dump (meca.cc:957-990): ddim.txt, diagonal.txt, drhs.bin, dsol.bin,
        dpts.bin and dmat.bin in `directory` (None: the working directory)"""
        d = None if directory is None else str(directory).encode()
        self._chk(self.L.cymg_dump(self.h, d))

    def matrix_entries(self, which):
        n = C.c_uint32()
        self._chk(self.L.cymg_get_matrix(self.h, which, C.byref(n), None, None, None))
        rows = np.empty(n.value, np.uint32)
        cols = np.empty(n.value, np.uint32)
        vals = np.empty(n.value)
        self._chk(self.L.cymg_get_matrix(self.h, which, C.byref(n), _p(rows), _p(cols), _p(vals)))
        return rows, cols, vals

    def enable_timing(self, on=True):
        self._chk(self.L.cymg_enable_timing(self.h, int(on)))

    def set_option(self, option, value):
        self._chk(self.L.cymg_set_option(self.h, option, value))

    def autotune(self, reps=20):
        """Pick the block-inverse method, the tile shape and the staging of
        the blocks by timing them on the current system (after setup);
        returns the choices."""
        t, b, ms = C.c_int(), C.c_int(), C.c_double()
        self._chk(self.L.cymg_autotune(self.h, int(reps), C.byref(t), C.byref(b), C.byref(ms)))
        return {"tile_points": t.value, "tile_threads": self.get_option(OPT_TILE_THREADS),
                "block_inverse": b.value, "block_stage": self.get_option(OPT_BLOCK_STAGE),
                "launch_ms": ms.value}

    def get_option(self, option):
        v = C.c_int()
        self._chk(self.L.cymg_get_option(self.h, option, C.byref(v)))
        return v.value

    def time_kernel(self, which, reps=50):
        ms = C.c_double()
        self._chk(self.L.cymg_time_kernel(self.h, which, reps, C.byref(ms)))
        return ms.value

    def timings(self):
        t = Timings()
        self._chk(self.L.cymg_get_timings(self.h, C.byref(t)))
        return t.as_dict()

    # learned initial guess (no counterpart in the reference) ------------------------
    def guess_init(self, seed=0):
        self._chk(self.L.cymg_guess_init(self.h, int(seed)))

    def guess_weights(self):
        w = np.empty(GUESS_WEIGHTS, np.float32)
        self._chk(self.L.cymg_guess_weights(self.h, 0, _p(w)))
        return w

    def set_guess_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == GUESS_WEIGHTS
        self._chk(self.L.cymg_guess_weights(self.h, 1, _p(w)))

    def guess_record(self, capacity):
        self._chk(self.L.cymg_guess_record(self.h, int(capacity)))

    def guess_samples(self):
        n = C.c_uint32()
        self._chk(self.L.cymg_guess_samples(self.h, C.byref(n)))
        return n.value

    def guess_train(self, steps, batch=256, lr=1e-3, seed=0):
        loss = C.c_double()
        self._chk(self.L.cymg_guess_train(self.h, int(steps), int(batch), float(lr), int(seed), C.byref(loss)))
        return loss.value

    def guess_eval(self, batch=8192, seed=99):
        """loss of the current weights on `batch` recorded samples (no training step)"""
        loss = C.c_double()
        self._chk(self.L.cymg_guess_eval(self.h, int(batch), int(seed), C.byref(loss)))
        return loss.value

    def guess_check(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        assert X.shape[1] == GUESS_INPUTS and T.shape == (X.shape[0], GUESS_OUTPUTS)
        g = np.empty(GUESS_WEIGHTS, np.float32)
        loss = C.c_double()
        self._chk(self.L.cymg_guess_check(self.h, _p(X), _p(T), X.shape[0], C.byref(loss), _p(g)))
        return loss.value, g

    def guess_features(self):
        nf = len(self.np)
        X = np.empty((nf, GUESS_INPUTS), np.float32)
        s = np.empty(nf, np.float32)
        self._chk(self.L.cymg_guess_features(self.h, _p(X), _p(s)))
        return X, s

    def guess_predict(self, precondition=PRECOND_FP64):
        """(learned guess x0, its base M^-1 b) of the current system."""
        n = self.size
        x0 = np.empty(n)
        base = np.empty(n)
        self._chk(self.L.cymg_guess_predict(self.h, precondition, _p(x0), _p(base)))
        return x0, base


def load_batch(meca, batch, couples_on_device=True):
    """Register a Batch of independent systems and run one step's assembly."""
    sc = batch.scene
    meca.clear()
    for f in range(sc.nfib):
        meca.add_fiber(sc.nb_points[f], sc.segmentation[f], sc.rigidity[f])
    meca.set_systems(batch.first_fiber)
    meca.set_params(**sc.params())
    meca.prepare(sc.points)
    assemble(meca, sc, couples_on_device, True)


def load_scene(meca, scene, couples_on_device=True, upload_couples=True):
    """Register the scene's fibers and run one step's assembly in the order of
    Simul::setInteractions (simul_solve.cc:199-245): Meca::clear/add/prepare,
    then Fiber::setInteractions (sphere confinement), then Couple::setInteractions.
    Couples are emitted on the device from the uploaded couple list, or, with
    couples_on_device=False, registered one by one through the host interLink."""
    meca.clear()
    for f in range(scene.nfib):
        meca.add_fiber(scene.nb_points[f], scene.segmentation[f], scene.rigidity[f])
    for k in range(getattr(scene, "nsol", 0)):
        meca.add_solid(scene.sol_radius[k], scene.solid_points(k), int(scene.sol_timer[k]))
    meca.set_params(**scene.params())
    meca.prepare(scene.points)
    assemble(meca, scene, couples_on_device, upload_couples)


def assemble(meca, scene, couples_on_device=True, upload_couples=True):
    """One step's interactions, after prepare()."""
    if scene.confine_radius > 0:
        meca.confine_sphere(scene.confine_radius, scene.confine_stiffness)
    if scene.nlinks:
        if couples_on_device:
            if upload_couples:
                meca.set_couples(scene.link_fiber_a, scene.link_abs_a, scene.link_fiber_b,
                                 scene.link_abs_b, scene.link_stiffness)
                if getattr(scene, "link_state", None) is not None:
                    meca.set_couple_state(scene.link_state)
            meca.couples_interactions()
        else:
            a1, a2, ca = scene.interpolate(scene.link_fiber_a, scene.link_abs_a)
            b1, b2, cb = scene.interpolate(scene.link_fiber_b, scene.link_abs_b)
            state = getattr(scene, "link_state", None)
            for k in range(scene.nlinks):
                if state is not None and (int(state[k]) & 3) != 3:
                    continue      # only bridging couples interact (couple.cc:87-92)
                meca.inter_link_ii(int(a1[k]), int(a2[k]), float(ca[k]), int(b1[k]), int(b2[k]),
                                   float(cb[k]), float(scene.link_stiffness[k]))
    # Aster::setInteractions (aster.cc:53-89), after the couples (simul_solve.cc:238-239)
    if getattr(scene, "naster_links", 0) and couples_on_device:
        if upload_couples:
            meca.set_aster_links(scene.aster_kind, scene.aster_s1, scene.aster_s2, scene.aster_fiber,
                                 scene.aster_abs, scene.aster_w)
        meca.aster_interactions()
        return
    for k in range(getattr(scene, "naster_links", 0)):
        if scene.aster_kind[k] == 0:
            meca.inter_link_ee(int(scene.aster_a1[k]), int(scene.aster_b[k]), float(scene.aster_w[k]))
        else:
            meca.inter_link_ie(int(scene.aster_a1[k]), int(scene.aster_a2[k]), float(scene.aster_ca[k]),
                               int(scene.aster_b[k]), float(scene.aster_w[k]))
