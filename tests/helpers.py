"""Shared helpers of the parity tests."""
import numpy as np

from elmm_amd import capi


def interior(a, n):
    """Slice of a ghosted (lb=-2) array covering indices 1..n per axis."""
    return a[3:3 + n[0], 3:3 + n[1], 3:3 + n[2]]


def component_extents(cfg):
    per = [cfg.btype[1] == capi.PERIODIC, cfg.btype[3] == capi.PERIODIC, cfg.btype[5] == capi.PERIODIC]
    ns = (cfg.nx, cfg.ny, cfg.nz)
    return [[ns[d] if (d != c or per[d]) else ns[d] - 1 for d in range(3)] for c in range(3)]


def rel_err(a, b):
    """max |a-b| / max |b| (0 when both vanish)."""
    den = float(np.max(np.abs(b)))
    num = float(np.max(np.abs(a - b)))
    if den == 0.0:
        return num
    return num / den


def compare_state(a, b, cfg, fields=None):
    """Relative max errors of the interior of each field between two solvers
    (a Solver or Oracle: anything with download)."""
    ext = component_extents(cfg)
    pr_n = (cfg.nx, cfg.ny, cfg.nz)
    out = {}
    names = fields or ["U", "V", "W", "Pr", "Visc"] + [f"S{s}" for s in range(cfg.nscalars)]
    for name in names:
        if name in ("U", "V", "W"):
            c = "UVW".index(name)
            fa, fb = a.download(c), b.download(c)
            out[name] = rel_err(interior(fa, ext[c]), interior(fb, ext[c]))
        elif name == "Pr":
            out[name] = rel_err(interior(a.download(capi.F_PR), pr_n), interior(b.download(capi.F_PR), pr_n))
        elif name == "Visc":
            out[name] = rel_err(interior(a.download(capi.F_VISC), pr_n), interior(b.download(capi.F_VISC), pr_n))
        elif name == "TDiff":
            out[name] = rel_err(interior(a.download(capi.F_TDIFF), pr_n), interior(b.download(capi.F_TDIFF), pr_n))
        elif name.startswith("S"):
            s = int(name[1:])
            out[name] = rel_err(interior(a.download(capi.F_SCALAR + s), pr_n),
                                interior(b.download(capi.F_SCALAR + s), pr_n))
    return out


def velocity_rel_l2(a, b, cfg):
    """Relative L2 error of the velocity vector (U, V, W interiors) of a
    against b: ||u_a - u_b||_2 / ||u_b||_2."""
    ext = component_extents(cfg)
    num = den = 0.0
    for c in range(3):
        fa = interior(a.download(c), ext[c])
        fb = interior(b.download(c), ext[c])
        num += float(np.sum((fa - fb) ** 2))
        den += float(np.sum(fb ** 2))
    return (num / den) ** 0.5
