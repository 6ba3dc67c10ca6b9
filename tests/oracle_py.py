"""ctypes wrapper of the CPU oracle (oracle/lib/libelmm_oracle.so).

Test infrastructure: used only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed baseline.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "lib", "libelmm_oracle.so")

_lib = None


def _load():
    global _lib
    if _lib is None:
        import sys
        sys.path.insert(0, os.path.join(ROOT, "gpu-elmm_amd"))
        from elmm_amd import capi
        lib = C.CDLL(ORACLE_LIB)
        pd = C.POINTER(C.c_double)
        lib.oracle_create.restype = C.c_void_p
        lib.oracle_create.argtypes = [C.POINTER(capi.Config), C.POINTER(capi.Lists)]
        lib.oracle_destroy.argtypes = [C.c_void_p]
        lib.oracle_upload.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int32), pd]
        lib.oracle_download.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int32), pd]
        lib.oracle_init.argtypes = [C.c_void_p]
        lib.oracle_step.argtypes = [C.c_void_p, C.c_int32, C.POINTER(capi.Diag)]
        lib.oracle_divergence.argtypes = [C.c_void_p]
        lib.oracle_divergence.restype = C.c_double
        lib.oracle_last_error.restype = C.c_char_p
        lib.oracle_op_bound_uvw.argtypes = [C.c_void_p]
        lib.oracle_op_sgs.argtypes = [C.c_void_p]
        lib.oracle_op_advection.argtypes = [C.c_void_p]
        lib.oracle_op_diffusion.argtypes = [C.c_void_p]
        lib.oracle_op_kappa.argtypes = [C.c_void_p, C.c_int32]
        lib.oracle_op_poisson.argtypes = [C.c_void_p, pd, pd]
        lib.oracle_op_residual.argtypes = [C.c_void_p, pd, pd]
        lib.oracle_op_residual.restype = C.c_double
        lib.oracle_det_log.argtypes = [C.c_double]
        lib.oracle_det_log.restype = C.c_double
        for name in ("oracle_det_exp", "oracle_det_atan", "oracle_psim_mo", "oracle_psim_mo_mod",
                     "oracle_psih_mo_mod"):
            getattr(lib, name).argtypes = [C.c_double]
            getattr(lib, name).restype = C.c_double
        lib.oracle_det_pow.argtypes = [C.c_double, C.c_double]
        lib.oracle_det_pow.restype = C.c_double
        lib.oracle_mo_flux_ustar.argtypes = [C.c_double] * 8
        lib.oracle_mo_flux_ustar.restype = C.c_double
        lib.oracle_mo_dirichlet.argtypes = [pd, pd] + [C.c_double] * 8
        lib.oracle_mo_dirichlet_moist.argtypes = [pd, pd, pd] + [C.c_double] * 11
        lib.oracle_op_wm_fluxes.argtypes = [C.c_void_p]
        lib.oracle_wm_list.argtypes = [C.c_void_p, C.c_int32, pd, pd]
        lib.oracle_wm_list.restype = C.c_int32
        for name in ("oracle_op_ib_momentum", "oracle_op_mass_sources", "oracle_op_null", "oracle_op_ib_visc",
                     "oracle_op_bound_pr", "oracle_op_bound_visc"):
            getattr(lib, name).argtypes = [C.c_void_p]
        lib.oracle_op_wm_div.argtypes = [C.c_void_p, pd]
        lib.oracle_op_wm_div.restype = C.c_int32
        lib.oracle_op_poisson_mg.argtypes = [C.c_void_p, pd, pd, pd]
        lib.oracle_op_poisson_mg.restype = C.c_int32
        lib.oracle_op_wm_visc.argtypes = [C.c_void_p]
        for name in ("oracle_op_buoyancy", "oracle_op_temp_fix", "oracle_op_bound_temperature",
                     "oracle_op_moist_fill", "oracle_op_bound_moisture"):
            getattr(lib, name).argtypes = [C.c_void_p]
        lib.oracle_wm_ustar.argtypes = [C.c_void_p, pd]
        lib.oracle_wm_ustar.restype = C.c_int32
        lib.oracle_op_scalar_extra.argtypes = [C.c_void_p, C.c_int32]
        lib.oracle_op_bound_scalar.argtypes = [C.c_void_p, C.c_int32]
        _lib = lib
    return _lib


F_USTAR, F_VSTAR, F_WSTAR, F_RHS, F_SADV = 8, 9, 10, 11, 32


def det_log(x):
    """The logarithm of the oracle's smooth-wall law."""
    return _load().oracle_det_log(float(x))


def det_fn(name, *args):
    """The oracle's fixed-operation exp, atan, pow and Monin-Obukhov
    functions (oracle_<name>)."""
    return getattr(_load(), "oracle_" + name)(*[float(a) for a in args])


def mo_dirichlet_moist(ustar, tflux, qflux, dist, z0, vel, temp, surf_temp, moist, surf_moist, nu, kappa, g, tref):
    """WM_MO_Dirichlet_ustar_tfl_mfl (z0H = z0): the updated (ustar, tflux,
    qflux)."""
    u, t, q = C.c_double(ustar), C.c_double(tflux), C.c_double(qflux)
    _load().oracle_mo_dirichlet_moist(C.byref(u), C.byref(t), C.byref(q), dist, z0, vel, temp, surf_temp, moist,
                                      surf_moist, nu, kappa, g, tref)
    return u.value, t.value, q.value


def mo_dirichlet(ustar, tflux, dist, z0, vel, tempdif, nu, kappa, g, tref):
    """WM_MO_Dirichlet_ustar_tfl (z0H = z0): the updated (ustar, tflux)."""
    u, t = C.c_double(ustar), C.c_double(tflux)
    _load().oracle_mo_dirichlet(C.byref(u), C.byref(t), dist, z0, vel, tempdif, nu, kappa, g, tref)
    return u.value, t.value


class Oracle:
    def __init__(self, cfg, case):
        lib = _load()
        self.cfg = cfg
        self.case = case
        self._h = lib.oracle_create(C.byref(cfg), C.byref(case.view))
        if not self._h:
            raise RuntimeError("oracle_create: " + lib.oracle_last_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            _lib.oracle_destroy(self._h)
            self._h = None

    __del__ = close

    def shape(self):
        return (self.cfg.nx + 6, self.cfg.ny + 6, self.cfg.nz + 6)

    def upload(self, field, arr, lb=-2):
        a = np.asfortranarray(arr, dtype=np.float64)
        ext = (C.c_int32 * 3)(*a.shape)
        if _lib.oracle_upload(self._h, field, lb, ext, a.ctypes.data_as(C.POINTER(C.c_double))):
            raise RuntimeError(_lib.oracle_last_error().decode())

    def download(self, field, lb=-2, shape=None):
        shape = shape or self.shape()
        a = np.zeros(shape, dtype=np.float64, order="F")
        ext = (C.c_int32 * 3)(*shape)
        if _lib.oracle_download(self._h, field, lb, ext, a.ctypes.data_as(C.POINTER(C.c_double))):
            raise RuntimeError(_lib.oracle_last_error().decode())
        return a

    def init(self):
        if _lib.oracle_init(self._h):
            raise RuntimeError(_lib.oracle_last_error().decode())

    def step(self, n=1):
        from elmm_amd import capi
        d = capi.Diag()
        rc = _lib.oracle_step(self._h, n, C.byref(d))
        if rc:
            raise RuntimeError(_lib.oracle_last_error().decode())
        return d

    def divergence(self):
        return _lib.oracle_divergence(self._h)

    def op_bound_uvw(self):
        _lib.oracle_op_bound_uvw(self._h)

    def op_sgs(self):
        _lib.oracle_op_sgs(self._h)

    def op_advection(self):
        _lib.oracle_op_advection(self._h)

    def op_diffusion(self):
        _lib.oracle_op_diffusion(self._h)

    def op_wm_fluxes(self):
        _lib.oracle_op_wm_fluxes(self._h)

    def wm_list(self, comp, direction):
        """(fluxp, fluxm) of the momentum wall-model list of component comp
        (0..2) and direction 0..5 (x-, x+, y-, y+, z-, z+)."""
        pd = C.POINTER(C.c_double)
        n = _lib.oracle_wm_list(self._h, comp * 6 + direction, None, None)
        fp, fm = np.zeros(n), np.zeros(n)
        _lib.oracle_wm_list(self._h, comp * 6 + direction, fp.ctypes.data_as(pd), fm.ctypes.data_as(pd))
        return fp, fm

    def op_ib_momentum(self):
        _lib.oracle_op_ib_momentum(self._h)

    def op_mass_sources(self):
        _lib.oracle_op_mass_sources(self._h)

    def op_null(self):
        _lib.oracle_op_null(self._h)

    def op_ib_visc(self):
        _lib.oracle_op_ib_visc(self._h)

    def op_wm_div(self):
        n = _lib.oracle_op_wm_div(self._h, None)
        out = np.zeros(n)
        _lib.oracle_op_wm_div(self._h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def op_wm_visc(self):
        _lib.oracle_op_wm_visc(self._h)

    def wm_ustar(self):
        """WMPoints%ustar in list order."""
        n = _lib.oracle_wm_ustar(self._h, None)
        out = np.zeros(n)
        _lib.oracle_wm_ustar(self._h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def op_buoyancy(self):
        _lib.oracle_op_buoyancy(self._h)

    def op_temp_fix(self):
        _lib.oracle_op_temp_fix(self._h)

    def op_bound_temperature(self):
        _lib.oracle_op_bound_temperature(self._h)

    def op_moist_fill(self):
        _lib.oracle_op_moist_fill(self._h)

    def op_bound_moisture(self):
        _lib.oracle_op_bound_moisture(self._h)

    def op_scalar_extra(self, sc=0):
        _lib.oracle_op_scalar_extra(self._h, sc)

    def op_bound_pr(self):
        _lib.oracle_op_bound_pr(self._h)

    def op_bound_scalar(self, sc=0):
        _lib.oracle_op_bound_scalar(self._h, sc)

    def op_bound_visc(self):
        _lib.oracle_op_bound_visc(self._h)

    def op_kappa(self, sc=0):
        _lib.oracle_op_kappa(self._h, sc)

    def op_poisson(self, rhs):
        """Direct solve of the FD2 Poisson problem; rhs, result: (nx,ny,nz) Fortran order."""
        r = np.asfortranarray(rhs, dtype=np.float64)
        out = np.zeros_like(r, order="F")
        pd = C.POINTER(C.c_double)
        _lib.oracle_op_poisson(self._h, r.ctypes.data_as(pd), out.ctypes.data_as(pd))
        return out

    def op_poisson_mg(self, rhs):
        """Multigrid solve of the same problem: (phi, V-cycles, final max
        residual)."""
        r = np.asfortranarray(rhs, dtype=np.float64)
        out = np.zeros_like(r, order="F")
        res = C.c_double()
        pd = C.POINTER(C.c_double)
        cyc = _lib.oracle_op_poisson_mg(self._h, r.ctypes.data_as(pd), out.ctypes.data_as(pd), C.byref(res))
        return out, cyc, res.value

    def op_residual(self, phi_ghosted, rhs_ghosted):
        pd = C.POINTER(C.c_double)
        p = np.asfortranarray(phi_ghosted, dtype=np.float64)
        r = np.asfortranarray(rhs_ghosted, dtype=np.float64)
        return _lib.oracle_op_residual(self._h, p.ctypes.data_as(pd), r.ctypes.data_as(pd))
