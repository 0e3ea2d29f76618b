"""NetCDF output and input of bg_flood_amd (the library's bgf_nc_* entry points; no libnetcdf).

BuqOutput writes BG_Flood's BUQ map output: per output variable and level a grid
<var>_P<l>(time, yy_P<l>, xx_P<l>) with every block's 16x16 interior at its place and the fill
value elsewhere, plus the block table (creatncfileBUQ / defncvarBUQ / writencvarstepBUQ,
src/Write_netcdf.cu:118-500, 560-840, 1215-1360), in the NetCDF classic format with 64-bit offsets
(the reference writes the same content into a NETCDF4 file).

read_var / read_grid read NetCDF classic files (versions 1, 2 and 5), e.g. bathymetry grids and
hot-start files (readfileinfo / readstaticforcing, src/Read_netcdf.cu, src/InitEvolv.cu:565-835).
"""
import ctypes

import numpy as np

from . import BgfError, _blocks_struct, _check, _param_struct, lib

FILL = np.float32(9.9692e36)


def _bind():
    L = lib()
    if getattr(L, "_ncio_bound", False):
        return L
    vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    L.bgf_nc_buq_create.argtypes = [ctypes.c_char_p, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                    ctypes.POINTER(vp)]
    L.bgf_nc_buq_write.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(vp)]
    L.bgf_nc_buq_close.argtypes = [vp]
    L.bgf_nc_inq.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ip, ctypes.POINTER(ctypes.c_longlong), ip]
    L.bgf_nc_dim_names.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    L.bgf_nc_get.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_longlong, dp]
    L.bgf_nc_get_att.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, dp, ip]
    for n in ("bgf_nc_buq_create", "bgf_nc_buq_write", "bgf_nc_buq_close", "bgf_nc_inq", "bgf_nc_dim_names",
              "bgf_nc_get", "bgf_nc_get_att"):
        getattr(L, n).restype = ctypes.c_int
    L._ncio_bound = True
    return L


def _b(s):
    return s.encode() if isinstance(s, str) else s


def zone_extent(param, blocks):
    """(xmax, ymax) of the domain: the far edge of the farthest active block (Param.xmax and
    Param.ymax of a domain rounded up to whole blocks, src/ReadInput.cu:1280-1282)."""
    act = np.asarray(blocks["active"])[:param["nblk"]]
    lev = np.asarray(blocks["level"])[act]
    d = param["dx"] * np.where(lev < 0, 2.0 ** (-lev), 1.0 / 2.0 ** np.maximum(lev, 0))
    xo = np.asarray(blocks["xo"])[act]
    yo = np.asarray(blocks["yo"])[act]
    return float(param["xo"] + np.max(xo + 15.5 * d)), float(param["yo"] + np.max(yo + 15.5 * d))


class BuqOutput:
    """A BG_Flood map output file: BuqOutput(path, param, blocks, outvars) creates it, write(time,
    fields) appends one output (fields: name -> array in BUQ layout, nblkmem*324 values), close()
    ends it.  outvars name the variables in order (Param.outvars); smallnc, scalefactor and
    addoffset as Param.smallnc / scalefactor / addoffset."""

    def __init__(self, path, param, blocks, outvars, xmax=None, ymax=None, smallnc=0, scalefactor=0.01,
                 addoffset=0.0):
        L = _bind()
        self.path = str(path)
        self.outvars = list(outvars)
        self.nblkmem = int(param["nblkmem"])
        self.precision = param["precision"]
        if xmax is None or ymax is None:
            ex, ey = zone_extent(param, blocks)
            xmax = ex if xmax is None else xmax
            ymax = ey if ymax is None else ymax
        self._keep = {}
        P = _param_struct(param)
        B = _blocks_struct(blocks, self.nblkmem, self._keep)
        names = (ctypes.c_char_p * max(1, len(self.outvars)))(*[_b(n) for n in self.outvars])
        h = ctypes.c_void_p()
        _check(L.bgf_nc_buq_create(_b(self.path), ctypes.byref(P), ctypes.byref(B), float(xmax), float(ymax),
                                   len(self.outvars), names, int(smallnc), float(scalefactor), float(addoffset),
                                   ctypes.byref(h)))
        self.h = h

    def write(self, time, fields):
        if self.h is None:
            raise BgfError("output file is closed")
        dbl = self.precision in ("fp64", 0)
        T = np.float64 if dbl else np.float32
        arrs = []
        for n in self.outvars:
            if n not in fields:
                raise BgfError("no values for output variable %r" % n)
            a = np.ascontiguousarray(np.asarray(fields[n], T).ravel())
            if a.size != self.nblkmem * 324:
                raise BgfError("%s: expected %d values in BUQ layout, got %d" % (n, self.nblkmem * 324, a.size))
            arrs.append(a)
        ptrs = (ctypes.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])
        _check(_bind().bgf_nc_buq_write(self.h, float(time), 0 if dbl else 1, ptrs))

    def close(self):
        if self.h is not None:
            _check(_bind().bgf_nc_buq_close(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def var_shape(path, name):
    """(shape, is_record) of a variable; a record dimension counts the records written."""
    L = _bind()
    nd = ctypes.c_int()
    shp = (ctypes.c_longlong * 8)()
    rec = ctypes.c_int()
    _check(L.bgf_nc_inq(_b(str(path)), _b(name), ctypes.byref(nd), shp, ctypes.byref(rec)))
    return tuple(int(shp[k]) for k in range(nd.value)), bool(rec.value)


def var_dims(path, name):
    buf = ctypes.create_string_buffer(4096)
    _check(_bind().bgf_nc_dim_names(_b(str(path)), _b(name), buf, 4096))
    s = buf.value.decode()
    return s.split(",") if s else []


def has_var(path, name):
    try:
        var_shape(path, name)
        return True
    except BgfError:
        return False


def read_var(path, name, rec=None):
    """A fixed variable, or record rec of a record variable (all records if rec is None), as
    float64 in its file shape."""
    L = _bind()
    shape, isrec = var_shape(path, name)
    if not isrec:
        out = np.empty(int(np.prod(shape)) if shape else 1)
        _check(L.bgf_nc_get(_b(str(path)), _b(name), -1, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out.reshape(shape)
    slab = shape[1:]
    recs = range(shape[0]) if rec is None else [rec]
    out = np.empty((len(recs),) + slab)
    for k, r in enumerate(recs):
        buf = np.empty(int(np.prod(slab)) if slab else 1)
        _check(L.bgf_nc_get(_b(str(path)), _b(name), int(r), buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        out[k] = buf.reshape(slab)
    return out if rec is None else out[0]


def read_att(path, att, var=None):
    L = _bind()
    n = ctypes.c_int(64)
    buf = (ctypes.c_double * 64)()
    _check(L.bgf_nc_get_att(_b(str(path)), _b(var or ""), _b(att), buf, ctypes.byref(n)))
    v = [buf[k] for k in range(min(n.value, 64))]
    return v[0] if len(v) == 1 else v


def read_grid(path, var, step=0):
    """A gridded variable as (x, y, z[ny, nx]): the last two dimensions are y and x and their
    coordinate variables carry the same names (readfileinfo, src/Read_netcdf.cu); a third,
    leading dimension (time) is indexed by step (readstaticforcing's step)."""
    dims = var_dims(path, var)
    if len(dims) not in (2, 3):
        raise BgfError("%s: expected 2 or 3 dimensions, got %s" % (var, dims))
    x = read_var(path, dims[-1])
    y = read_var(path, dims[-2])
    shape, isrec = var_shape(path, var)
    if len(dims) == 3:
        if isrec:
            z = read_var(path, var, rec=step)
        else:
            z = read_var(path, var)[step]
    else:
        z = read_var(path, var)
    return np.asarray(x, np.float64), np.asarray(y, np.float64), np.asarray(z)


def check_var_name(path, *names):
    """First of names present in the file, "" if none (checkncvarname, src/Read_netcdf.cu)."""
    for n in names:
        if has_var(path, n):
            return n
    return ""


def _raster(path, name, step):
    """A grid variable as a float32 raster with increasing x and y: (values, xo, yo, dx)."""
    x, y, z = read_grid(path, name, step)
    z = np.asarray(z, np.float32)
    if len(x) < 2 or len(y) < 2:
        raise BgfError("%s: a grid needs at least 2 x 2 points" % name)
    if x[1] < x[0]:
        x, z = x[::-1], z[:, ::-1]
    if y[1] < y[0]:
        y, z = y[::-1], z[::-1, :]
    dx = float(x[1] - x[0])
    return np.ascontiguousarray(z), float(x[0]), float(y[0]), dx


HOTSTART_NAMES = {  # checkncvarname candidates (src/InitEvolv.cu:636-640)
    "zb": ("zb", "z", "ZB", "Z", "zb_P0"),
    "zs": ("zs", "eta", "ZS", "ETA", "zs_P0"),
    "h": ("h", "hh", "hhh", "hhhh", "h_P0"),
    "u": ("u", "uu", "uvel", "UVEL", "u_P0"),
    "v": ("v", "vv", "vvel", "VVEL", "v_P0"),
}


def hotstart(domain, path, step=0, zb=None):
    """State of a hot start from a NetCDF file (readhotstartfile, src/InitEvolv.cu:598-835): zb, zs,
    h, u and v are looked up under the reference's names, read at `step` and sampled onto the
    domain's mesh (interp2BUQ); zs is raised to the bed, h is max(zs - zb, 0) when the file has zs
    only, zs = zb + h when it has h only; missing u, v are zero.  zb: the domain's bed (BUQ array),
    used when the file has none.  Returns dict(zb, h, u, v) in BUQ layout (the domain's T), or None
    when the file has neither zs nor h (the reference then falls back to a cold start).  A file with
    both zs and h gives h from the file (the state here is zb, h, u, v)."""
    p = domain.param
    T = np.float64 if p["precision"] == "fp64" else np.float32
    S = p["nblkmem"] * 324
    names = {k: check_var_name(path, *v) for k, v in HOTSTART_NAMES.items()}
    if not names["zs"] and not names["h"]:
        return None

    def sample(k):
        vals, xo, yo, dx = _raster(path, names[k], step)
        return domain.sample_raster(vals, xo, yo, dx)

    if names["zb"]:
        zbv = sample("zb")
    elif zb is not None:
        zbv = np.asarray(zb, T).copy()
    else:
        raise BgfError("%s has no bed (zb) and none was given" % path)
    zs = None
    if names["zs"]:
        zs = np.maximum(sample("zs"), zbv).astype(T)  # utils::max(zs, zb)
    if names["h"]:
        h = sample("h")
    else:
        h = np.maximum(zs - zbv, T(0.0)).astype(T)
    u = sample("u") if names["u"] else np.zeros(S, T)
    v = sample("v") if names["v"] else np.zeros(S, T)
    from . import interior_mask
    m = interior_mask(p["nblkmem"], np.asarray(domain.blocks["active"])[:p["nblk"]])
    for a in (h, u, v):
        a[~m] = 0
    return dict(zb=zbv, h=h, u=u, v=v)
