"""Shared helpers of the tests: build the oracle and the GPU domain from one scenario."""
import numpy as np

from oracle import pyoracle as po
import bg_flood_amd as bgf

ORACLE_KEYS = ("g", "eps", "CFL", "theta", "dx", "delta", "xo", "yo", "lat", "mask", "frictionmodel", "aoibnd",
               "ForceMassConserve", "wetdryfix", "nblk", "nblkmem", "minlevel", "maxlevel")


def oracle_for(sc, precision=None, nthreads=4):
    p = sc["param"]
    prec = precision or p["precision"]
    kw = {k: p[k] for k in ORACLE_KEYS}
    kw.update(spherical=p.get("spherical", 0), Radius=p.get("Radius", 6371220.0))
    kw.update(cf=p.get("cf", 0.0), nx=sc["nbx"] * 16, ny=sc["nby"] * 16, rain_on=p["rain_on"], rain=p["rain"],
              doubleprecision=1 if prec == "fp64" else 0, nthreads=nthreads, dtinit=p.get("dtinit", -1.0))
    op = po.make_param(**kw)
    o = po.Oracle(op, sc["blocks"])
    T = np.float64 if prec == "fp64" else np.float32
    o.set_state(*[np.asarray(sc[k]).astype(T) for k in ("zb", "h", "u", "v", "cf")])
    return o


def domain_for(sc, precision=None):
    p = dict(sc["param"])
    if precision:
        p["precision"] = precision
    d = bgf.Domain(p, sc["blocks"])
    T = d.dtype
    d.set_state(*[np.asarray(sc[k]).astype(T) for k in ("zb", "h", "u", "v", "cf")])
    return d


def interior(sc):
    return bgf.interior_mask(sc["param"]["nblkmem"], sc["blocks"]["active"][:sc["param"]["nblk"]])


def orc_param_with(op, **kw):
    """A copy of an oracle parameter block with some fields changed (nblk, nblkmem after adaptation)."""
    q = po.make_param(**{f: getattr(op, f) for f, _ in po.OrcParam._fields_})
    for k, v in kw.items():
        setattr(q, k, v)
    return q
