"""ctypes binding of the CPU oracle (oracle/meca_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and as the timed CPU restatement.
The product path never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libmeca_oracle.so")

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [("time_step", C.c_double), ("kT", C.c_double), ("viscosity", C.c_double),
                ("hydro_radius0", C.c_double), ("hydro_radius1", C.c_double)]


class Motor(C.Structure):
    _fields_ = [("max_speed", C.c_double), ("stall_force", C.c_double), ("limit_speed", C.c_int),
                ("hold_growing_end", C.c_int)]


class SolveParams(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("tol_mode", C.c_int),
                ("precondition", C.c_int), ("max_iter", C.c_int)]


class Report(C.Structure):
    _fields_ = [("iterations", C.c_int), ("flag", C.c_int), ("converged", C.c_int),
                ("fallback", C.c_int), ("precond_used", C.c_int), ("residual", C.c_double),
                ("threshold", C.c_double), ("noise_level", C.c_double),
                ("total_iterations", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


class StericParams(C.Structure):
    _fields_ = [("extension", C.c_double * 3), ("max_range", C.c_double), ("stiff_push", C.c_double),
                ("stiff_pull", C.c_double)]


class StericCall(C.Structure):
    _fields_ = [("kind", C.c_int), ("p", C.c_uint32 * 4), ("ca", C.c_double), ("cb", C.c_double),
                ("len", C.c_double), ("w", C.c_double)]


class FrameRecord(C.Structure):
    _fields_ = [("number", C.c_uint32), ("mark", C.c_int32), ("prop", C.c_uint32), ("signature", C.c_uint32),
                ("cut_wished", C.c_double), ("abscissa", C.c_double)]


def frame_records(records):
    """ctypes array of FrameRecord from (number, mark, prop, signature, cut_wished, abscissa) tuples"""
    arr = (FrameRecord * max(1, len(records)))()
    for k, r in enumerate(records):
        arr[k] = FrameRecord(*r)
    return arr


def build():
    """Compile the oracle with its Makefile (gcc only)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "orc_create": (vp, [C.c_int]),
        "orc_destroy": (None, [vp]),
        "orc_set_params": (None, [vp, C.POINTER(Params)]),
        "orc_clear": (None, [vp]),
        "orc_add_fiber": (C.c_int, [vp, C.c_uint32, C.c_double, C.c_double]),
        "orc_inter_couples": (None, [vp, C.c_uint32, vp, vp, vp, vp, vp, vp]),
        "orc_add_fiber_ex": (C.c_int, [vp, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double]),
        "orc_add_solid": (C.c_int, [vp, C.c_uint32, _f64p, _f64p, C.c_int]),
        "orc_get_solid": (None, [vp, C.c_int, _f64p]),
        "orc_solid_timer": (C.c_int, [vp, C.c_int]),
        "orc_nb_points": (C.c_uint32, [vp]),
        "orc_mat_index": (C.c_uint32, [vp, C.c_int]),
        "orc_best_nb_points": (C.c_uint32, [C.c_double]),
        "orc_prepare": (C.c_int, [vp, _f64p]),
        "orc_interpolate": (None, [vp, C.c_int, C.c_double, C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
        "orc_add_pure_force": (None, [vp, C.c_uint32, _f64p]),
        "orc_inter_link_ee": (None, [vp, C.c_uint32, C.c_uint32, C.c_double]),
        "orc_inter_link_ie": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_double]),
        "orc_inter_link_ii": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32,
                                     C.c_uint32, C.c_double, C.c_double]),
        "orc_inter_long_link_ee": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_double]),
        "orc_inter_clamp_e": (None, [vp, C.c_uint32, _f64p, C.c_double]),
        "orc_inter_clamp_i": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, _f64p, C.c_double]),
        "orc_inter_long_clamp": (None, [vp, C.c_uint32, _f64p, C.c_double, C.c_double]),
        "orc_inter_plane_e": (None, [vp, C.c_uint32, _f64p, _f64p, C.c_double]),
        "orc_confine_sphere": (None, [vp, C.c_double, C.c_double]),
        "orc_steric": (C.c_int, [vp, C.POINTER(StericParams), vp, _f64p, _f64p, C.c_int, vp, C.c_uint32,
                                 C.POINTER(C.c_uint32)]),
        "orc_write_frame": (C.c_size_t, [vp, vp, C.c_double, C.c_char_p, C.c_int, C.c_char_p, C.c_size_t,
                                         C.c_char_p, C.c_size_t, vp, C.c_size_t]),
        "orc_add_pure_force_i": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, _f64p]),
        "orc_inter_long_link_ie": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_double, C.c_double]),
        "orc_inter_long_link_ii": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double]),
        "orc_inter_side_link_ie": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_double, C.c_double]),
        "orc_inter_side_link_ii": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double]),
        "orc_inter_side_side_link": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double]),
        "orc_inter_sliding_link_ie": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_double]),
        "orc_inter_sliding_link_ii": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32, C.c_double, C.c_double]),
        "orc_inter_side_sliding_link_ie": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_double, C.c_double]),
        "orc_inter_side_sliding_link_ii": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double]),
        "orc_inter_torque": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_double]),
        "orc_inter_side_clamp": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, _f64p, C.c_double, C.c_double]),
        "orc_inter_sliding_clamp": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, _f64p, C.c_double]),
        "orc_inter_plane_i": (None, [vp, C.c_uint32, C.c_uint32, C.c_double, _f64p, _f64p, C.c_double]),
        "orc_step_couples": (None, [vp, C.c_uint32, vp, _f64p, vp, _f64p, _f64p, vp, C.POINTER(Motor)]),
        "orc_nearby_segments": (C.c_uint32, [vp, _f64p, C.c_double, C.c_int, C.c_int, C.c_uint32,
                                             vp, vp, vp, vp]),
        "orc_setup": (C.c_int, [vp, vp]),
        "orc_iterate": (C.c_int, [vp, C.POINTER(SolveParams), vp, C.POINTER(Report)]),
        "orc_apply": (C.c_int, [vp]),
        "orc_solve": (C.c_int, [vp, C.POINTER(SolveParams), vp, C.POINTER(Report)]),
        "orc_multiply": (None, [vp, _f64p, _f64p]),
        "orc_compute_preconditioner": (C.c_int, [vp]),
        "orc_precondition": (None, [vp, _f64p, _f64p]),
        "orc_get_block": (None, [vp, C.c_int, _f64p]),
        "orc_get_inverse_block": (C.c_int, [vp, C.c_int, _f64p]),
        "orc_project": (None, [vp, C.c_int, _f64p, _f64p, C.c_double]),
        "orc_add_rigidity": (None, [vp, C.c_int, _f64p, _f64p]),
        "orc_add_projection_diff": (None, [vp, C.c_int, _f64p, _f64p]),
        "orc_compute_forces": (None, [vp]),
        "orc_forces_linear": (None, [vp, _f64p, _f64p, C.c_int]),
        "orc_get_vector": (None, [vp, C.c_int, _f64p]),
        "orc_get_fiber_points": (None, [vp, _f64p]),
        "orc_get_tensions": (None, [vp, _f64p]),
        "orc_get_mobility": (C.c_double, [vp, C.c_int]),
        "orc_get_jjt": (None, [vp, C.c_int, _f64p, _f64p]),
        "orc_get_matrix_nr": (C.c_uint32, [vp, C.c_int, vp, vp]),
        "orc_get_matrix_dense": (None, [vp, C.c_int, _f64p]),
        "orc_dgetrf": (C.c_int, [C.c_int, _f64p, C.c_int, _i32p]),
        "orc_dgetri": (C.c_int, [C.c_int, _f64p, C.c_int, _i32p]),
        "orc_dpttrf": (C.c_int, [C.c_int, _f64p, _f64p]),
        "orc_dptts2": (None, [C.c_int, _f64p, _f64p, _f64p]),
        "orc_dgtsv": (C.c_int, [C.c_int, _f64p, _f64p, _f64p, _f64p]),
        "orc_reshape": (C.c_int, [C.c_int, C.c_uint32, _f64p, _f64p, C.c_double]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


PTS, SOL, BAS, RHS, FOR = range(5)


class OracleMeca:
    """Thin object wrapper over one oracle Meca instance."""

    def __init__(self, dim, params=None):
        self.L = lib()
        self.dim = dim
        self.h = self.L.orc_create(dim)
        if not self.h:
            raise ValueError("dim must be 2 or 3")
        self.nfib = 0
        self.np = []
        self.nseg = 0
        if params is not None:
            self.set_params(**params)

    def __del__(self):
        try:
            if self.h:
                self.L.orc_destroy(self.h)
        except Exception:
            pass

    def set_params(self, time_step, kT, viscosity, hydro_radius0=0.0125, hydro_radius1=10.0):
        p = Params(time_step, kT, viscosity, hydro_radius0, hydro_radius1)
        self.L.orc_set_params(self.h, C.byref(p))

    def clear(self):
        self.L.orc_clear(self.h)
        self.nfib = 0
        self.np = []
        self.nseg = 0

    def add_fiber(self, nb_points, segmentation, rigidity, hydro_radius0=0.0, hydro_radius1=0.0,
                  surface_height=0.0):
        """A fiber; hydro radii 0 take the Meca's parameters, surface_height > 0
        selects the surface drag model (fiber.cc:401-447)."""
        r = self.L.orc_add_fiber_ex(self.h, int(nb_points), float(segmentation), float(rigidity),
                                    float(hydro_radius0), float(hydro_radius1), float(surface_height))
        if r < 0:
            raise ValueError("fiber needs at least 2 points")
        self.nfib += 1
        self.np.append(int(nb_points))
        self.nseg += int(nb_points) - 1
        return r

    def add_solid(self, radius, points, reshape_timer=0):
        """A Solid Mecable (after the fibers) with per-point radii and its
        reference shape fixed from `points` (DIM per point)."""
        r = np.ascontiguousarray(radius, dtype=np.float64)
        pts = np.ascontiguousarray(points, dtype=np.float64).ravel()
        assert pts.size == self.dim * r.size
        i = self.L.orc_add_solid(self.h, r.size, r, pts, int(reshape_timer))
        if i < 0:
            raise ValueError("invalid solid")
        self.np.append(int(r.size))
        return i

    def solid(self, i):
        out = np.empty(15)
        self.L.orc_get_solid(self.h, i, out)
        return {"drag": out[0], "drag_rot": out[1], "mom2d": out[2], "center": out[3:6].copy(),
                "mom": out[6:15].reshape(3, 3, order="F").copy()}

    def solid_timer(self, i):
        return int(self.L.orc_solid_timer(self.h, i))

    @property
    def nb_points(self):
        return int(self.L.orc_nb_points(self.h))

    @property
    def size(self):
        return self.dim * self.nb_points

    def mat_index(self, f):
        return int(self.L.orc_mat_index(self.h, f))

    def prepare(self, points):
        pts = np.ascontiguousarray(points, dtype=np.float64).ravel()
        assert pts.size == self.size
        r = self.L.orc_prepare(self.h, pts)
        if r == -3:
            raise RuntimeError("zero drag in Solid, or Solid momentum matrix not factorisable")
        if r:
            raise RuntimeError("could not build Fiber projection")

    def interpolate(self, fiber, abscissa):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_double()
        self.L.orc_interpolate(self.h, fiber, abscissa, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def inter_link_ii(self, a1, a2, ca, b1, b2, cb, w):
        self.L.orc_inter_link_ii(self.h, a1, a2, ca, b1, b2, cb, w)

    def inter_couples(self, fiber_a, abs_a, fiber_b, abs_b, stiffness, state=None):
        fa = np.ascontiguousarray(fiber_a, dtype=np.uint32)
        fb = np.ascontiguousarray(fiber_b, dtype=np.uint32)
        aa = np.ascontiguousarray(abs_a, dtype=np.float64)
        ab = np.ascontiguousarray(abs_b, dtype=np.float64)
        w = np.ascontiguousarray(stiffness, dtype=np.float64)
        st = None if state is None else np.ascontiguousarray(state, dtype=np.uint8)
        self.L.orc_inter_couples(self.h, fa.size, fa.ctypes.data, aa.ctypes.data, fb.ctypes.data, ab.ctypes.data,
                                 w.ctypes.data, None if st is None else st.ctypes.data)

    def inter_link_ie(self, a1, a2, ca, b, w):
        self.L.orc_inter_link_ie(self.h, a1, a2, ca, b, w)

    def inter_link_ee(self, a, b, w):
        self.L.orc_inter_link_ee(self.h, a, b, w)

    def inter_long_link_ee(self, a, b, length, w):
        self.L.orc_inter_long_link_ee(self.h, a, b, length, w)

    def inter_clamp_e(self, a, g, w):
        self.L.orc_inter_clamp_e(self.h, a, np.ascontiguousarray(g, dtype=np.float64), w)

    def inter_clamp_i(self, a1, a2, ca, g, w):
        self.L.orc_inter_clamp_i(self.h, a1, a2, ca, np.ascontiguousarray(g, dtype=np.float64), w)

    def inter_long_clamp(self, a, center, length, w):
        self.L.orc_inter_long_clamp(self.h, a, np.ascontiguousarray(center, dtype=np.float64), length, w)

    def inter_plane_e(self, a, direction, g, w):
        self.L.orc_inter_plane_e(self.h, a, np.ascontiguousarray(direction, dtype=np.float64),
                                 np.ascontiguousarray(g, dtype=np.float64), w)

    def add_pure_force(self, a, force):
        self.L.orc_add_pure_force(self.h, a, np.ascontiguousarray(force, dtype=np.float64))

    def add_pure_force_i(self, a1, a2, ca, force):
        self.L.orc_add_pure_force_i(self.h, a1, a2, ca, np.ascontiguousarray(force, dtype=np.float64))

    def inter_long_link_ie(self, a1, a2, ca, b, length, w):
        self.L.orc_inter_long_link_ie(self.h, a1, a2, ca, b, length, w)

    def inter_long_link_ii(self, a1, a2, ca, b1, b2, cb, length, w):
        self.L.orc_inter_long_link_ii(self.h, a1, a2, ca, b1, b2, cb, length, w)

    def inter_side_link_ie(self, a1, a2, ca, b, length, w):
        self.L.orc_inter_side_link_ie(self.h, a1, a2, ca, b, length, w)

    def inter_side_link_ii(self, a1, a2, ca, b1, b2, cb, length, w):
        self.L.orc_inter_side_link_ii(self.h, a1, a2, ca, b1, b2, cb, length, w)

    def inter_side_side_link(self, a1, a2, ca, b1, b2, cb, length, w):
        if self.L.orc_inter_side_side_link(self.h, a1, a2, ca, b1, b2, cb, length, w):
            raise ValueError("interSideSideLink() was not implemented in 3D")

    def inter_sliding_link_ie(self, a1, a2, ca, b, w):
        self.L.orc_inter_sliding_link_ie(self.h, a1, a2, ca, b, w)

    def inter_sliding_link_ii(self, a1, a2, ca, b1, b2, cb, w):
        self.L.orc_inter_sliding_link_ii(self.h, a1, a2, ca, b1, b2, cb, w)

    def inter_side_sliding_link_ie(self, a1, a2, ca, b, length, w):
        self.L.orc_inter_side_sliding_link_ie(self.h, a1, a2, ca, b, length, w)

    def inter_side_sliding_link_ii(self, a1, a2, ca, b1, b2, cb, length, w):
        self.L.orc_inter_side_sliding_link_ii(self.h, a1, a2, ca, b1, b2, cb, length, w)

    def inter_torque(self, a1, a2, ca, b1, b2, cb, cosinus, sinus, w):
        if self.L.orc_inter_torque(self.h, a1, a2, ca, b1, b2, cb, cosinus, sinus, w):
            raise ValueError("interTorque2D exists in 2D only")

    def inter_side_clamp(self, a1, a2, ca, g, length, w):
        self.L.orc_inter_side_clamp(self.h, a1, a2, ca, np.ascontiguousarray(g, dtype=np.float64), length, w)

    def inter_sliding_clamp(self, a1, a2, ca, g, w):
        self.L.orc_inter_sliding_clamp(self.h, a1, a2, ca, np.ascontiguousarray(g, dtype=np.float64), w)

    def inter_plane_i(self, a1, a2, ca, direction, g, w):
        self.L.orc_inter_plane_i(self.h, a1, a2, ca, np.ascontiguousarray(direction, dtype=np.float64),
                                 np.ascontiguousarray(g, dtype=np.float64), w)

    def write_frame(self, records, time, date, real_bytes=4, before=b"", after=b""):
        """Simul::writeObjects (simul_file.cc:402-489) in binary mode: the frame's bytes;
        records: per Mecable (number, mark, prop, signature, cut_wished, abscissa)"""
        arr = frame_records(records)
        d = date.encode() if isinstance(date, str) else date
        n = self.L.orc_write_frame(self.h, arr, time, d, real_bytes, before, len(before), after, len(after),
                                   None, 0)
        buf = C.create_string_buffer(n)
        self.L.orc_write_frame(self.h, arr, time, d, real_bytes, before, len(before), after, len(after), buf, n)
        return buf.raw[:n]

    def steric(self, extension, max_range, push, pull, steric, radius, range_, brute_force=False, cap=1 << 20):
        """Simul::setStericInteractions (simul_solve.cc:139-188) with PointGrid; returns
        the list of Meca calls made: (kind, points, ca, cb, len, weight) with kind 0
        interLongLink PE-PE, 1 interSideSlidingLink PI-PE, 2 PI-PI"""
        sp = StericParams()
        for d in range(3):
            sp.extension[d] = extension[d] if d < len(extension) else 0.0
        sp.max_range, sp.stiff_push, sp.stiff_pull = max_range, push, pull
        st = np.ascontiguousarray(steric, dtype=np.int32)
        rd = np.ascontiguousarray(radius, dtype=np.float64)
        rg = np.ascontiguousarray(range_, dtype=np.float64)
        log = (StericCall * cap)()
        n = C.c_uint32()
        rc = self.L.orc_steric(self.h, C.byref(sp), st.ctypes.data, rd, rg, int(brute_force), log, cap, C.byref(n))
        if rc == -3:
            raise ValueError("simul:steric_max_range is too short")
        if rc:
            raise ValueError(f"steric grid error {rc}")
        assert n.value <= cap
        return [(c.kind, tuple(c.p), c.ca, c.cb, c.len, c.w) for c in log[:n.value]]

    def confine_sphere(self, radius, stiffness):
        self.L.orc_confine_sphere(self.h, radius, stiffness)

    def step_couples(self, fiber_a, abs_a, fiber_b, abs_b, stiffness, state, max_speed, stall_force,
                     limit_speed=True, hold_growing_end=False):
        """Couple::stepAA/AF/FA with motor hands (couple.cc:125-150); abs_a,
        abs_b (float64) and state (uint8) are updated in place."""
        fa = np.ascontiguousarray(fiber_a, dtype=np.uint32)
        fb = np.ascontiguousarray(fiber_b, dtype=np.uint32)
        w = np.ascontiguousarray(stiffness, dtype=np.float64)
        for arr, t in ((abs_a, np.float64), (abs_b, np.float64), (state, np.uint8)):
            assert arr.dtype == t and arr.flags.c_contiguous
        mp = Motor(float(max_speed), float(stall_force), int(bool(limit_speed)), int(bool(hold_growing_end)))
        self.L.orc_step_couples(self.h, fa.size, fa.ctypes.data, abs_a, fb.ctypes.data, abs_b, w,
                                state.ctypes.data, C.byref(mp))

    def nearby_segments(self, place, range_, inclusive=False, exclude=-1):
        """FiberGrid::nearbySegments by brute force (fiber_grid.cc:302-337):
        (fiber, point, abscissa, dist_sqr) of the segments near `place`, in
        fiber order."""
        w = np.ascontiguousarray(place, dtype=np.float64)
        n = self.L.orc_nearby_segments(self.h, w, float(range_), int(bool(inclusive)), int(exclude), 0,
                                       None, None, None, None)
        f = np.zeros(n, np.uint32)
        p = np.zeros(n, np.uint32)
        a = np.zeros(n, np.float64)
        d = np.zeros(n, np.float64)
        self.L.orc_nearby_segments(self.h, w, float(range_), int(bool(inclusive)), int(exclude), n,
                                   f.ctypes.data, p.ctypes.data, a.ctypes.data, d.ctypes.data)
        return f, p, a, d

    def setup(self, gauss=None):
        g = None if gauss is None else np.ascontiguousarray(gauss, dtype=np.float64)
        r = self.L.orc_setup(self.h, _ptr(g))
        if r:
            raise RuntimeError("setup before prepare")

    def iterate(self, tolerance, tol_mode=0, precondition=1, max_iter=0, guess=None):
        sp = SolveParams(tolerance, tol_mode, precondition, max_iter)
        rep = Report()
        g = None if guess is None else np.ascontiguousarray(guess, dtype=np.float64)
        r = self.L.orc_iterate(self.h, C.byref(sp), _ptr(g), C.byref(rep))
        out = rep.as_dict()
        out["failed"] = bool(r)
        return out

    def apply(self):
        if self.L.orc_apply(self.h):
            raise RuntimeError("Cannot rescale the Solid, because it has collapsed!")

    def solve(self, tolerance, gauss=None, tol_mode=0, precondition=1, max_iter=0):
        self.setup(gauss)
        rep = self.iterate(tolerance, tol_mode, precondition, max_iter)
        if rep["failed"]:
            raise RuntimeError("convergence failure in solver")
        self.apply()
        return rep

    def multiply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self.L.orc_multiply(self.h, x, y)
        return y

    def compute_preconditioner(self):
        self.L.orc_compute_preconditioner(self.h)

    def precondition(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self.L.orc_precondition(self.h, x, y)
        return y

    def get_block(self, f):
        bs = self.dim * self.np[f]
        out = np.zeros(bs * bs)
        self.L.orc_get_block(self.h, f, out)
        return out.reshape(bs, bs).T.copy()     # column-major -> row-major matrix

    def get_inverse_block(self, f):
        bs = self.dim * self.np[f]
        out = np.zeros(bs * bs)
        ok = self.L.orc_get_inverse_block(self.h, f, out)
        return out.reshape(bs, bs).T.copy() if ok else None

    def project(self, f, x, sc=1.0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self.L.orc_project(self.h, f, x, y, sc)
        return y

    def add_rigidity(self, f, x, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, dtype=np.float64).copy()
        self.L.orc_add_rigidity(self.h, f, x, y)
        return y

    def add_projection_diff(self, f, x, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, dtype=np.float64).copy()
        self.L.orc_add_projection_diff(self.h, f, x, y)
        return y

    def forces_linear(self, x, with_rigidity=True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self.L.orc_forces_linear(self.h, x, y, int(with_rigidity))
        return y

    def compute_forces(self):
        self.L.orc_compute_forces(self.h)

    def vector(self, which):
        out = np.empty(self.size)
        self.L.orc_get_vector(self.h, which, out)
        return out

    def fiber_points(self):
        out = np.empty(self.size)
        self.L.orc_get_fiber_points(self.h, out)
        return out

    def tensions(self):
        out = np.empty(self.nseg)
        self.L.orc_get_tensions(self.h, out)
        return out

    def mobility(self, f):
        return self.L.orc_get_mobility(self.h, f)

    def jjt(self, f):
        S = self.np[f] - 1
        d = np.empty(S)
        e = np.empty(max(S - 1, 1))
        self.L.orc_get_jjt(self.h, f, d, e)
        return d, e[:S - 1]

    def matrix_nr(self, which):
        nbe = self.L.orc_get_matrix_nr(self.h, which, None, None)
        ija = np.empty(nbe, dtype=np.uint32)
        sa = np.empty(nbe)
        self.L.orc_get_matrix_nr(self.h, which, ija.ctypes.data_as(C.c_void_p),
                                 sa.ctypes.data_as(C.c_void_p))
        return ija, sa

    def matrix_dense(self, which):
        n = self.nb_points if which == 0 else self.size
        out = np.empty(n * n)
        self.L.orc_get_matrix_dense(self.h, which, out)
        return out.reshape(n, n)


# LAPACK restatements -------------------------------------------------------

def dgetrf(a):
    a = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    n = a.shape[0]
    buf = a.ravel(order="F").copy()
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().orc_dgetrf(n, buf, n, ipiv)
    return buf.reshape(n, n, order="F"), ipiv, info


def dgetri(lu, ipiv):
    n = lu.shape[0]
    buf = np.asfortranarray(lu, dtype=np.float64).ravel(order="F").copy()
    info = lib().orc_dgetri(n, buf, n, np.ascontiguousarray(ipiv, dtype=np.int32))
    return buf.reshape(n, n, order="F"), info


def dpttrf(d, e):
    d = np.array(d, dtype=np.float64)
    e = np.array(e, dtype=np.float64) if len(e) else np.zeros(1)
    info = lib().orc_dpttrf(len(d), d, e)
    return d, e[:len(d) - 1], info


def dptts2(d, e, b):
    b = np.array(b, dtype=np.float64)
    e = np.array(e, dtype=np.float64) if len(e) else np.zeros(1)
    lib().orc_dptts2(len(d), np.ascontiguousarray(d, dtype=np.float64), e, b)
    return b


def dgtsv(dl, d, du, b):
    dl = np.array(dl, dtype=np.float64) if len(dl) else np.zeros(1)
    du = np.array(du, dtype=np.float64) if len(du) else np.zeros(1)
    d = np.array(d, dtype=np.float64)
    b = np.array(b, dtype=np.float64)
    info = lib().orc_dgtsv(len(d), dl, d, du, b)
    return b, info


def reshape(dim, src, cut):
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.empty_like(src)
    ns = src.size // dim - 1
    failed = lib().orc_reshape(dim, ns, src, dst, cut)
    return dst, failed


def best_nb_points(ratio):
    return int(lib().orc_best_nb_points(float(ratio)))


def assemble_scene(m, sc):
    """Records the scene's interactions on an OracleMeca in the call order of
    Simul::setInteractions (simul_solve.cc:215-242): fibers' sphere
    confinement, the crosslinkers (Couple::setInteractions -> Meca::interLink),
    then the asters (Aster::setInteractions, aster.cc:53-89)."""
    if sc.confine_radius > 0:
        m.confine_sphere(sc.confine_radius, sc.confine_stiffness)
    if sc.nlinks:
        # only bridging couples interact (couple.cc:87-92); one C call
        m.inter_couples(sc.link_fiber_a, sc.link_abs_a, sc.link_fiber_b, sc.link_abs_b, sc.link_stiffness,
                        getattr(sc, "link_state", None))
    for k in range(getattr(sc, "naster_links", 0)):
        if sc.aster_kind[k] == 0:
            m.inter_link_ee(int(sc.aster_a1[k]), int(sc.aster_b[k]), float(sc.aster_w[k]))
        else:
            m.inter_link_ie(int(sc.aster_a1[k]), int(sc.aster_a2[k]), float(sc.aster_ca[k]),
                            int(sc.aster_b[k]), float(sc.aster_w[k]))


def register_scene(m, sc):
    """Meca::add of the scene's Mecables: fibers, then Solids."""
    for f in range(sc.nfib):
        m.add_fiber(sc.nb_points[f], sc.segmentation[f], sc.rigidity[f])
    for k in range(getattr(sc, "nsol", 0)):
        m.add_solid(sc.sol_radius[k], sc.solid_points(k), int(sc.sol_timer[k]))


def from_scene(sc):
    """OracleMeca loaded with one step of a scene: registration (Meca::add),
    prepare, then the interactions."""
    m = OracleMeca(sc.dim, sc.params())
    register_scene(m, sc)
    m.prepare(sc.points)
    assemble_scene(m, sc)
    return m
