"""ctypes binding of the C ABI in include/elmm_gpu.h.

This is the Python host-side mirror of the reference's time-step interface:
``Solver.step`` plays the role of ``TMarchRK3`` (src/tsteps.f90:24-243),
``Solver.init`` the initial-condition tail (src/initial.f90:4284-4342) and
``Case`` the list set-up of InitBoundaryConditions (src/initial.f90:
4507-4681).  Arrays cross the boundary in Fortran order with the lower bound
-2 in every direction (the reference's U/V/W/Scalar allocation).
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libelmm_hip.so")

# sides and boundary codes (src/parameters.f90:233, 260-261)
WE, EA, SO, NO, BO, TO = range(6)
NOSLIP, FREESLIP, PERIODIC, DIRICHLET, NEUMANN, CONSTFLUX = 1, 2, 3, 4, 5, 6
WM_NEUTRAL, WM_SURFACE_TEMP, WM_SURFACE_FLUX = 0, 1, 2
FP64, FP32_REFINED = 0, 1
POISSON_FFT, POISSON_MG = 2, 3
F_U, F_V, F_W, F_PR, F_VISC, F_TDIFF, F_PHI, F_Q, F_MOISTURE, F_TEMPERATURE, F_SCALAR = 0, 1, 2, 3, 4, 5, 6, 7, 14, 15, 16


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
        ("stretched_z", C.c_int32),
        ("zw", C.POINTER(C.c_double)),
        ("btype", C.c_int32 * 6),
        ("z0_wall", C.c_double * 6),
        ("scal_btype", C.c_int32 * 6),
        ("side_scal", C.c_double * 6),
        ("nu", C.c_double), ("nu_scalar", C.c_double), ("c_smag", C.c_double),
        ("filter_ratio", C.c_double), ("dt", C.c_double),
        ("pr_gradient_x", C.c_double), ("pr_gradient_y", C.c_double),
        ("advection_method", C.c_int32), ("sgs_model", C.c_int32), ("wall_model", C.c_int32),
        ("nscalars", C.c_int32), ("poisson_solver", C.c_int32),
        ("poisson_tol", C.c_double),
        ("cfl_check", C.c_int32),
        ("cfl_max", C.c_double),
        ("precision", C.c_int32), ("ib_mass_sources", C.c_int32),
        ("refine_iters", C.c_int32),
        ("variable_dt", C.c_int32), ("delta_diag", C.c_int32),
        ("cfl_target", C.c_double), ("dt_max", C.c_double), ("dt_min", C.c_double),
        ("start_time", C.c_double), ("end_time", C.c_double),
        ("autotune", C.c_int32), ("mg_cycles_max", C.c_int32),
        ("enable_buoyancy", C.c_int32), ("grav_acc", C.c_double), ("temperature_ref", C.c_double),
        ("temp_btype", C.c_int32 * 6), ("side_temp", C.c_double * 6),
        ("enable_moisture", C.c_int32), ("moist_btype", C.c_int32 * 6), ("moist_pad", C.c_int32),
        ("moisture_ref", C.c_double), ("side_moist", C.c_double * 6),
        ("reserved", C.c_int32 * 2),
    ]


_pi32 = C.POINTER(C.c_int32)
_pu8 = C.POINTER(C.c_uint8)
_pd = C.POINTER(C.c_double)


class Lists(C.Structure):
    _fields_ = [
        ("prtype", _pi32), ("utype", _pi32), ("vtype", _pi32), ("wtype", _pi32),
        ("scfl_mask", _pu8 * 3),
        ("wm_n", C.c_int32 * 18), ("wm_ijk", _pi32 * 18), ("wm_dist", _pd * 18), ("wm_z0", _pd * 18),
        ("wmp_n", C.c_int32), ("wmp_ijk", _pi32), ("wmp_dist", _pd), ("wmp_z0", _pd),
        ("scfl_n", C.c_int32 * 3), ("scfl_ijk", _pi32 * 3),
        ("ib_n", C.c_int32 * 4), ("ib_ijk", _pi32 * 4), ("ib_ptr", _pi32 * 4),
        ("ib_src_ijk", _pi32 * 4), ("ib_coef", _pd * 4),
        ("null_n", C.c_int32 * 3), ("null_ijk", _pi32 * 3),
        ("src_n", C.c_int32), ("src_ijk", _pi32), ("src_scalar", _pi32), ("src_flux", _pd),
        ("wm_tclass", C.POINTER(C.c_int8) * 18), ("wm_tval", _pd * 18),
        ("wmp_tclass", C.POINTER(C.c_int8)), ("wmp_tval", _pd),
        ("wm_qval", _pd * 18), ("wm_qflux", _pd * 18), ("wmp_qval", _pd), ("wmp_qflux", _pd),
    ]


class Diag(C.Structure):
    _fields_ = [("delta", C.c_double), ("cfl", C.c_double), ("max_div", C.c_double),
                ("poisson_res", C.c_double), ("refinements", C.c_int32), ("steps_done", C.c_int32),
                ("stored_div", C.c_double), ("dt", C.c_double), ("time", C.c_double),
                ("reserved", C.c_int32 * 4)]


T_NGROUPS = 16


class Request(C.Structure):
    _fields_ = [("cfg", C.c_void_p), ("c", C.c_void_p), ("nsteps", C.c_int32), ("pad", C.c_int32),
                ("U", C.c_void_p), ("V", C.c_void_p), ("W", C.c_void_p), ("Pr", C.c_void_p),
                ("scalars", C.c_void_p),
                ("out_U", C.c_void_p), ("out_V", C.c_void_p), ("out_W", C.c_void_p), ("out_Pr", C.c_void_p),
                ("out_scalars", C.c_void_p)]


class RequestStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("steps_done", C.c_int32), ("pad", C.c_int32),
                ("latency_s", C.c_double), ("max_div", C.c_double)]


class Timers(C.Structure):
    _fields_ = [("ms", C.c_double * T_NGROUPS), ("launches", C.c_int64 * T_NGROUPS)]


class Tiles(C.Structure):
    _fields_ = [("mom_tx", C.c_int32), ("mom_ty", C.c_int32), ("mom_zchunk", C.c_int32),
                ("sgs_zchunk", C.c_int32), ("scalar_zchunk", C.c_int32), ("tuned", C.c_int32),
                ("reserved", C.c_int32 * 2)]


class Cuboid(C.Structure):
    _fields_ = [("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
                ("ymax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double), ("z0", C.c_double),
                ("temperature_flux", C.c_double), ("moisture_flux", C.c_double)]


class PointSource(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("flux", C.c_double),
                ("scalar", C.c_int32), ("pad", C.c_int32)]


# Every symbol the header declares, with its ctypes signature.
EXPORTS = {
    "elmm_case_build": (C.c_int, [C.POINTER(Config), C.c_int32, C.POINTER(Cuboid), C.c_int32,
                                  C.POINTER(PointSource), C.POINTER(C.c_void_p)]),
    "elmm_case_build_device": (C.c_int, [C.POINTER(Config), C.c_int32, C.POINTER(Cuboid), C.c_int32,
                                         C.POINTER(PointSource), C.POINTER(C.c_void_p)]),
    "elmm_case_view": (C.c_int, [C.c_void_p, C.POINTER(Lists)]),
    "elmm_case_free": (None, [C.c_void_p]),
    "elmm_create": (C.c_int, [C.POINTER(Config), C.POINTER(Lists), C.POINTER(C.c_void_p)]),
    "elmm_create_case": (C.c_int, [C.POINTER(Config), C.c_void_p, C.POINTER(C.c_void_p)]),
    "elmm_copy_probe": (C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "elmm_query": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "elmm_footprint": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_int64)]),
    "elmm_case_footprint": (C.c_int, [C.POINTER(Config), C.c_void_p, C.POINTER(C.c_int64)]),
    "elmm_server_create": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "elmm_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_int64)]),
    "elmm_server_progress": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "elmm_server_status": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(RequestStatus)]),
    "elmm_server_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "elmm_server_destroy": (C.c_int, [C.c_void_p]),
    "elmm_upload": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int32), _pd]),
    "elmm_download": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int32), _pd]),
    "elmm_init": (C.c_int, [C.c_void_p]),
    "elmm_step": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Diag)]),
    "elmm_step_async": (C.c_int, [C.c_void_p, C.c_int32]),
    "elmm_wait": (C.c_int, [C.c_void_p, C.POINTER(Diag)]),
    "elmm_divergence": (C.c_int, [C.c_void_p, _pd]),
    "elmm_divergence_noalloc": (C.c_int, [C.c_void_p, _pd]),
    "elmm_wm_ustar": (C.c_int, [C.c_void_p, C.c_int32, _pd, C.POINTER(C.c_int32)]),
    "elmm_destroy": (C.c_int, [C.c_void_p]),
    "elmm_vtk_write": (C.c_int, [C.c_char_p, C.POINTER(Config), _pd, _pd, _pd, _pd, _pd, C.c_int32,
                                 C.POINTER(_pd)]),
    "elmm_vtk_read": (C.c_int, [C.c_char_p, C.POINTER(Config), _pd, _pd, _pd, _pd, C.c_int32, C.POINTER(_pd)]),
    "elmm_frame_async": (C.c_int, [C.c_void_p, C.c_char_p]),
    "elmm_frame_wait": (C.c_int, [C.c_void_p]),
    "elmm_timers_enable": (C.c_int, [C.c_void_p, C.c_int32]),
    "elmm_get_tiles": (C.c_int, [C.c_void_p, C.POINTER(Tiles)]),
    "elmm_timers_read": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "elmm_timer_name": (C.c_char_p, [C.c_int32]),
    "elmm_version": (C.c_char_p, []),
    "elmm_last_error": (C.c_char_p, []),
}

_lib = None


class ElmmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"elmm error {code}: {msg}")
        self.code = code


def load(path=LIB_PATH):
    """Load libelmm_hip.so; raises if the HIP library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} not found: build the HIP library first (python __graft_entry__.py build)")
    lib = C.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise ElmmError(rc, _lib.elmm_last_error().decode())


def make_config(**kw):
    c = Config()
    c.filter_ratio = 1.0
    c.c_smag = 0.122
    c.advection_method = 4
    c.sgs_model = 1
    c.wall_model = 1
    c.poisson_solver = POISSON_FFT
    c.poisson_tol = 1e-10
    c.ib_mass_sources = 1
    zw = kw.pop("zw", None)
    for k, v in kw.items():
        if isinstance(getattr(c, k), C.Array):
            arr = getattr(c, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(c, k, v)
    if zw is not None:
        zw = np.ascontiguousarray(zw, dtype=np.float64)
        c._zw_keep = zw
        c.zw = zw.ctypes.data_as(_pd)
        c.stretched_z = 1
    return c


def with_precision(cfg, precision, refine_iters=0):
    """Copy of a configuration with another precision form."""
    c = Config()
    C.memmove(C.byref(c), C.byref(cfg), C.sizeof(Config))
    if hasattr(cfg, "_zw_keep"):
        c._zw_keep = cfg._zw_keep
    c.precision = precision
    c.refine_iters = refine_iters
    return c


def _np(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Case:
    """Masks and point lists of one case (elmm_case_build; device=True:
    elmm_case_build_device, the solid-body part built on the current GPU)."""

    def __init__(self, cfg, cuboids=(), sources=(), device=False):
        lib = load()
        self.cfg = cfg
        cub = (Cuboid * max(1, len(cuboids)))(*[Cuboid(*c) for c in cuboids])
        src = (PointSource * max(1, len(sources)))(
            *[PointSource(s[0], s[1], s[2], s[3], int(s[4]) if len(s) > 4 else 0, 0) for s in sources])
        h = C.c_void_p()
        build = lib.elmm_case_build_device if device else lib.elmm_case_build
        _check(build(C.byref(cfg), len(cuboids), cub, len(sources), src, C.byref(h)))
        self._h = h
        self.view = Lists()
        _check(lib.elmm_case_view(h, C.byref(self.view)))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.elmm_case_free(self._h)
            self._h = None

    def shape(self):
        return (self.cfg.nx + 6, self.cfg.ny + 6, self.cfg.nz + 6)

    def masks(self):
        """Prtype, Utype, Vtype, Wtype as (nx+6, ny+6, nz+6) Fortran arrays."""
        n = int(np.prod(self.shape()))
        out = []
        for name in ("prtype", "utype", "vtype", "wtype"):
            a = _np(getattr(self.view, name), n, np.int32)
            out.append(a.reshape(self.shape(), order="F"))
        return out

    def wm_list(self, l):
        v = self.view
        n = v.wm_n[l]
        return (_np(v.wm_ijk[l], 3 * n, np.int32).reshape(n, 3), _np(v.wm_dist[l], 3 * n, np.float64).reshape(n, 3),
                _np(v.wm_z0[l], n, np.float64))

    def wm_thermal(self, l):
        """Thermal class and value of the points of momentum list l (-1: the
        scalar wall-model points)."""
        v = self.view
        n = v.wmp_n if l < 0 else v.wm_n[l]
        tc = v.wmp_tclass if l < 0 else v.wm_tclass[l]
        tv = v.wmp_tval if l < 0 else v.wm_tval[l]
        if not tc:
            return np.zeros(n, np.int8), np.zeros(n)
        return _np(tc, n, np.int8), _np(tv, n, np.float64)

    def wm_moist(self, l):
        """Surface moisture and initial moisture flux of the points of list l
        (-1: the scalar wall-model points)."""
        v = self.view
        n = v.wmp_n if l < 0 else v.wm_n[l]
        qv = v.wmp_qval if l < 0 else v.wm_qval[l]
        qf = v.wmp_qflux if l < 0 else v.wm_qflux[l]
        if not qv:
            return np.zeros(n), np.zeros(n)
        return _np(qv, n, np.float64), _np(qf, n, np.float64)

    def wmp_list(self):
        v = self.view
        n = v.wmp_n
        return (_np(v.wmp_ijk, 3 * n, np.int32).reshape(n, 3), _np(v.wmp_dist, 3 * n, np.float64).reshape(n, 3),
                _np(v.wmp_z0, n, np.float64))

    def ib_list(self, c):
        v = self.view
        n = v.ib_n[c]
        ptr = _np(v.ib_ptr[c], n + 1, np.int32) if n else np.zeros(1, np.int32)
        nt = int(ptr[-1])
        return (_np(v.ib_ijk[c], 3 * n, np.int32).reshape(n, 3), ptr,
                _np(v.ib_src_ijk[c], 3 * nt, np.int32).reshape(nt, 3), _np(v.ib_coef[c], nt, np.float64))

    def null_list(self, c):
        v = self.view
        n = v.null_n[c]
        return _np(v.null_ijk[c], 3 * n, np.int32).reshape(n, 3)

    def counts(self):
        v = self.view
        return {"wm": [v.wm_n[l] for l in range(18)], "wmp": v.wmp_n, "scfl": [v.scfl_n[d] for d in range(3)],
                "ib": [v.ib_n[c] for c in range(4)], "null": [v.null_n[c] for c in range(3)], "src": v.src_n}


def ghosted_shape(cfg):
    return (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)


class Solver:
    """Device-resident LES case (elmm_create ... elmm_destroy)."""

    def __init__(self, cfg, case):
        lib = load()
        self.cfg = cfg
        self.case = case
        h = C.c_void_p()
        _check(lib.elmm_create(C.byref(cfg), C.byref(case.view), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.elmm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def upload(self, field, arr, lb=-2):
        a = np.asfortranarray(arr, dtype=np.float64)
        ext = (C.c_int32 * 3)(*a.shape)
        _check(_lib.elmm_upload(self._h, field, lb, ext, a.ctypes.data_as(_pd)))

    def download(self, field, lb=-2, shape=None):
        shape = shape or ghosted_shape(self.cfg)
        a = np.zeros(shape, dtype=np.float64, order="F")
        ext = (C.c_int32 * 3)(*shape)
        _check(_lib.elmm_download(self._h, field, lb, ext, a.ctypes.data_as(_pd)))
        return a

    def init(self):
        _check(_lib.elmm_init(self._h))

    def step(self, n=1, diag=True):
        d = Diag()
        _check(_lib.elmm_step(self._h, n, C.byref(d) if diag else None))
        return d

    def step_async(self, n=1):
        _check(_lib.elmm_step_async(self._h, n))

    def wait(self):
        _check(_lib.elmm_wait(self._h, None))

    def wait_diag(self):
        """elmm_wait with diagnostics (CFL, delta, dt, time, joined steps)."""
        d = Diag()
        _check(_lib.elmm_wait(self._h, C.byref(d)))
        return d

    def timers(self, enable=None, reset=False):
        """Per kernel group: (name, milliseconds, launches)."""
        if enable is not None:
            _check(_lib.elmm_timers_enable(self._h, 1 if enable else 0))
            return None
        t = Timers()
        _check(_lib.elmm_timers_read(self._h, C.byref(t), 1 if reset else 0))
        return {_lib.elmm_timer_name(g).decode(): (t.ms[g], t.launches[g]) for g in range(T_NGROUPS)
                if t.launches[g] > 0}

    def frame_async(self, directory):
        """elmm_frame_async: VTK files of the current state, written while
        later steps run."""
        _check(_lib.elmm_frame_async(self._h, os.fsencode(directory)))

    def frame_wait(self):
        _check(_lib.elmm_frame_wait(self._h))

    def tiles(self):
        """Tile shape and z chunks in use (elmm_get_tiles) as a dict."""
        t = Tiles()
        _check(_lib.elmm_get_tiles(self._h, C.byref(t)))
        return {f[0]: getattr(t, f[0]) for f in Tiles._fields_ if f[0] != "reserved"}

    def divergence(self):
        x = C.c_double()
        _check(_lib.elmm_divergence(self._h, C.byref(x)))
        return x.value

    def wm_ustar(self):
        """WMPoints%ustar of every scalar wall-model point (elmm_wm_ustar)."""
        n = C.c_int32()
        _check(_lib.elmm_wm_ustar(self._h, 0, None, C.byref(n)))
        out = np.zeros(max(n.value, 1))
        _check(_lib.elmm_wm_ustar(self._h, n.value, out.ctypes.data_as(_pd), C.byref(n)))
        return out[:n.value]

    def divergence_noalloc(self):
        """elmm_divergence_noalloc: the same maximum without device
        allocation (its partials in the library's static device array)."""
        x = C.c_double()
        _check(_lib.elmm_divergence_noalloc(self._h, C.byref(x)))
        return x.value


def footprint(cfg, case=None):
    """Device bytes of one case: elmm_footprint, or elmm_case_footprint
    (with the estimate of the case's surface lists) when a Case is given."""
    b = C.c_int64()
    if case is None:
        _check(load().elmm_footprint(C.byref(cfg), C.byref(b)))
    else:
        _check(load().elmm_case_footprint(C.byref(cfg), case._h, C.byref(b)))
    return b.value


class Server:
    """Continuous-batching request loop (elmm_server_*).  submit() copies the
    initial fields; the Case and the output arrays are kept referenced here
    until the request finishes."""

    def __init__(self, hbm_budget, max_active=64, chunk_steps=10):
        lib = load()
        h = C.c_void_p()
        _check(lib.elmm_server_create(int(hbm_budget), int(max_active), int(chunk_steps), C.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, cfg, case, fields, nsteps):
        U, V, W, Pr, scal = [np.asfortranarray(a, dtype=np.float64) for a in fields[:4]] + [fields[4]]
        scal = [np.asfortranarray(a, dtype=np.float64) for a in scal]
        shape = U.shape
        outs = [np.zeros(shape, order="F") for _ in range(4)]
        souts = [np.zeros(shape, order="F") for _ in range(cfg.nscalars)]
        sin = (C.c_void_p * max(1, len(scal)))(*[a.ctypes.data for a in scal]) if scal else None
        sout = (C.c_void_p * max(1, len(souts)))(*[a.ctypes.data for a in souts]) if souts else None
        r = Request(C.cast(C.pointer(cfg), C.c_void_p), case._h, int(nsteps), 0,
                    U.ctypes.data, V.ctypes.data, W.ctypes.data, Pr.ctypes.data,
                    C.cast(sin, C.c_void_p) if sin is not None else None,
                    outs[0].ctypes.data, outs[1].ctypes.data, outs[2].ctypes.data, outs[3].ctypes.data,
                    C.cast(sout, C.c_void_p) if sout is not None else None)
        rid = C.c_int64()
        _check(_lib.elmm_server_submit(self._h, C.byref(r), C.byref(rid)))
        self._keep[rid.value] = (cfg, case, outs, souts, sout)
        return rid.value

    def progress(self):
        a, q = C.c_int32(), C.c_int32()
        _check(_lib.elmm_server_progress(self._h, C.byref(a), C.byref(q)))
        return a.value, q.value

    def status(self, rid):
        st = RequestStatus()
        _check(_lib.elmm_server_status(self._h, int(rid), C.byref(st)))
        return st

    def result(self, rid):
        """(U, V, W, Pr, [scalars]) of a finished request (ghosted, lb -2)."""
        _, _, outs, souts, _ = self._keep[rid]
        return outs[0], outs[1], outs[2], outs[3], souts

    def release(self, rid):
        self._keep.pop(rid, None)

    def stats(self):
        cs, used = C.c_int64(), C.c_int64()
        _check(_lib.elmm_server_stats(self._h, C.byref(cs), C.byref(used)))
        return cs.value, used.value

    def close(self):
        if getattr(self, "_h", None):
            _lib.elmm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def copy_probe(nbytes=1 << 30, reps=5):
    """Achieved bandwidth (GB/s, read + write) of a plain device copy."""
    g = C.c_double()
    _check(load().elmm_copy_probe(int(nbytes), int(reps), C.byref(g)))
    return g.value


def vtk_write(directory, cfg, U, V, W, Pr, scalars=(), visc=None):
    """elmm_vtk_write: the reference's out.vtk, scalars.vtk and U/V/W.vtk
    from ghosted host arrays (lower bound -2)."""
    lib = load()
    arrs = [np.asfortranarray(a, dtype=np.float64) for a in (U, V, W, Pr)]
    sc = [np.asfortranarray(a, dtype=np.float64) for a in scalars]
    vs = np.asfortranarray(visc, dtype=np.float64) if visc is not None else None
    sp = (_pd * max(1, len(sc)))(*[a.ctypes.data_as(_pd) for a in sc])
    _check(lib.elmm_vtk_write(os.fsencode(directory), C.byref(cfg), *[a.ctypes.data_as(_pd) for a in arrs],
                              vs.ctypes.data_as(_pd) if vs is not None else None, len(sc), sp))


def vtk_read(directory, cfg, nscalars=0):
    """elmm_vtk_read: (U, V, W, Pr, [scalars]) as ghosted arrays."""
    lib = load()
    shape = ghosted_shape(cfg)
    out = [np.zeros(shape, order="F") for _ in range(4)]
    sc = [np.zeros(shape, order="F") for _ in range(nscalars)]
    sp = (_pd * max(1, len(sc)))(*[a.ctypes.data_as(_pd) for a in sc])
    _check(lib.elmm_vtk_read(os.fsencode(directory), C.byref(cfg), *[a.ctypes.data_as(_pd) for a in out],
                             len(sc), sp))
    return out[0], out[1], out[2], out[3], sc
