"""BG_param.txt reading for the parameters of the timestep path.

A line "key = value" sets a parameter (findparameter, src/ReadInput.cu:1689-1734):
  - the line is split at every '=' with empty pieces dropped (split, :1762-1775); fewer than two
    pieces means no parameter;
  - the key is the first piece with spaces trimmed (trim, :1891-1901, spaces only), compared with the
    parameter's name and aliases ignoring case (case_insensitive_compare, :1914-1921);
  - the value is the rest joined back with '=', cut at the first non-empty piece between '#'s, then
    at the first non-empty piece between ';'s, and space-trimmed; an empty value leaves the parameter
    unchanged.
Every line is offered to every parameter, so a key given twice keeps its last value.  Numbers are
parsed as std::stod / std::stoi do (the longest numeric prefix after leading whitespace; no number
is an error), booleans as readparambool (:2075-2091): 1/true/yes/on and -1/false/no/off, anything
else keeps the current value.

read_bg_param returns the Param fields the step reads, plus the run control and the mesh levels,
with the sanity rules of checkparamsanity (:1189-1440) that touch them: levels ordered and
initlevel clamped, zsoffset NaN -> 0, engine 5 -> CFL >= 0.25, an unset endtime -> 1/tiny, and
delta = dx (dx * Radius * pi / 180 on spherical grids, :1293-1305).  Forcing, boundary and output
keys are outside this path and ignored.

Deliberate differences, each for a reason of this path:
  - engine: the reference maps b/butt/buttinger/1 to engine 1 and every other value to engine 5
    (:193-213), so its Kurganov engine (2) cannot be chosen from the file; here "2" and "kurganov"
    give engine 2, the only engine this library provides.
  - ForceMassConserve: the reference also stores the value it read into totaltime (:437-447); here
    it sets ForceMassConserve only.
"""
import math
import re

PI_REF = 3.14159265  # pi of src/General.h:6, used for delta

# name -> (aliases, kind) for the parameters read here (src/ReadInput.cu:162-1188); kinds:
# f = std::stod, i = std::stoi, b = readparambool, engine = the engine rule, cf = a number unless the
# value has a letter (then it names a roughness map, :280-288), t = seconds (readinputtimetxt,
# src/utctime.cu:341-366; its date-time form needs the reference time and is not read here).
# lat is not read from the file (Param.lat stays 0, src/Param.h:38).
KEYS = {
    "g": (("g",), "f"),
    "eps": (("eps",), "f"),
    "CFL": (("CFL",), "f"),
    "theta": (("theta",), "f"),
    "dx": (("dx",), "f"),
    "xo": (("xo", "xmin"), "f"),
    "yo": (("yo", "ymin"), "f"),
    "mask": (("mask",), "f"),
    "cf": (("cf", "roughness", "cfmap"), "cf"),
    "frictionmodel": (("frictionmodel",), "i"),
    "aoibnd": (("aoibnd", "remainderbnd", "remainbndtype", "aoibndtype"), "i"),
    "ForceMassConserve": (("ForceMassConserve", "MassConservation", "MassCon", "forcemassconservation",
                           "forcevolumeconservation", "Volumeconservation", "VolumeCon", "ForceVolConserve"), "b"),
    "wetdryfix": (("wetdryfix", "reminstab", "fixinstab"), "b"),
    "spherical": (("spherical", "geo"), "b"),
    "Radius": (("Radius",), "f"),
    "minlevel": (("minlevel",), "i"),
    "maxlevel": (("maxlevel",), "i"),
    "initlevel": (("initlevel",), "i"),
    "doubleprecision": (("doubleprecision",), "i"),
    "dtinit": (("dtinit",), "f"),
    "zsinit": (("zsinit", "initzs"), "f"),
    "zsoffset": (("zsoffset",), "f"),
    "endtime": (("endtime", "stoptime", "end", "stop", "end_time", "stop_time"), "t"),
    "totaltime": (("totaltime", "inittime", "starttime", "start_time", "init_time", "start", "init"), "t"),
    "outputtimestep": (("outputtimestep", "outtimestep", "outputstep"), "f"),
    "engine": (("engine",), "engine"),
}

# src/Param.h:15-110 defaults of those parameters
DEFAULTS = dict(g=9.81, eps=0.0001, CFL=0.5, theta=1.3, dx=math.nan, xo=math.nan, yo=math.nan, lat=0.0,
                mask=9999.0, cf=0.0001, frictionmodel=0, aoibnd=0, ForceMassConserve=False, wetdryfix=True,
                spherical=False, Radius=6371220.0, minlevel=-99999, maxlevel=-99999, initlevel=0,
                doubleprecision=0, dtinit=-1.0, zsinit=math.nan, zsoffset=math.nan, endtime=1.7976931348623157e308,
                totaltime=0.0, outputtimestep=0.0, engine=1)

TRUE = ("1", "true", "yes", "on")
FALSE = ("-1", "false", "no", "off")

_FLOAT = re.compile(r"[\t\n\v\f\r ]*([+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan))", re.I)
_INT = re.compile(r"[\t\n\v\f\r ]*([+-]?\d+)")


def _split(s, delim):
    return [t for t in s.split(delim) if t]


def _trim(s):
    return s.strip(" ")


def find_parameter(aliases, line):
    """Value of the parameter named by aliases on this line, "" if the line does not set it."""
    pieces = _split(line, "=")
    if len(pieces) < 2:
        return ""
    left = _trim(pieces[0]).lower()
    if left not in [a.lower() for a in aliases]:
        return ""
    right = "=".join(pieces[1:])
    nohash = _split(right, "#")
    if not nohash:
        return ""
    first = _split(nohash[0], ";")
    return _trim(first[0]) if first else ""


def stod(v):
    m = _FLOAT.match(v)
    if not m:
        raise ValueError("not a number: %r" % v)
    return float(m.group(1))


def stoi(v):
    m = _INT.match(v)
    if not m:
        raise ValueError("not an integer: %r" % v)
    return int(m.group(1))


def read_param_bool(v, default):
    lv = v.lower()
    out = default
    if lv in TRUE:
        out = True
    if lv in FALSE:
        out = False
    return out


def _engine(v):
    lv = v.lower()
    if lv in ("b", "butt", "buttinger", "1"):
        return 1
    if lv in ("2", "kurganov"):
        return 2
    return 5


def parse_lines(text):
    """The parameters a BG_param text sets, before any sanity rule (Param fields)."""
    out = dict(DEFAULTS)
    for line in text.splitlines():
        for name, (aliases, kind) in KEYS.items():
            v = find_parameter(aliases, line)
            if not v:
                continue
            if kind == "f":
                out[name] = stod(v)
            elif kind == "i":
                out[name] = stoi(v)
            elif kind == "b":
                out[name] = read_param_bool(v, out[name])
            elif kind == "cf":
                if not any(c.isalpha() for c in v):
                    out[name] = stod(v)
            elif kind == "t":
                if len(_split(v, "T")) > 1:
                    raise ValueError("%s = %s: date-time values are not read here; give seconds" % (name, v))
                out[name] = stod(v)
            else:
                out[name] = _engine(v)
    return out


def sanity(p):
    """checkparamsanity's rules for these parameters (src/ReadInput.cu:1189-1440)."""
    p = dict(p)
    tiny = 0.0000001
    minlev, maxlev = p["minlevel"], p["maxlevel"]
    if minlev == -99999:
        minlev = p["initlevel"]
    if maxlev == -99999:
        maxlev = p["initlevel"]
    p["maxlevel"] = max(maxlev, minlev)
    p["minlevel"] = min(maxlev, minlev)
    p["initlevel"] = min(max(p["minlevel"], p["initlevel"]), p["maxlevel"])
    if math.isnan(p["zsoffset"]):
        p["zsoffset"] = 0.0
    if not math.isnan(p["dx"]):
        p["delta"] = p["dx"] * p["Radius"] * PI_REF / 180.0 if p["spherical"] else p["dx"]
    else:
        p["delta"] = math.nan  # the reference takes dx from the bathymetry file
    if p["engine"] == 5:
        p["CFL"] = max(p["CFL"], 0.25)
    if abs(p["endtime"] - DEFAULTS["endtime"]) <= tiny:
        p["endtime"] = 1.0 / tiny
    p["inittime"] = p["totaltime"]
    return p


def read_bg_param(text=None, path=None):
    """Param fields of a BG_param.txt (its text, or the file at path) for the timestep path."""
    if text is None:
        with open(path) as f:
            text = f.read()
    return sanity(parse_lines(text))


def domain_param(bgp, nblk, nblkmem, precision=None, rain_mmh=0.0):
    """bg_flood_amd.Domain parameters from read_bg_param's result and a mesh of nblk active blocks
    (levels minlevel..maxlevel).  precision defaults to the file's: fp64 if doubleprecision, else fp32."""
    from . import default_param

    for k in ("dx", "delta"):
        if math.isnan(bgp[k]):
            raise ValueError("dx is not set (the reference takes it from the bathymetry file): give dx")
    if precision is None:
        precision = "fp64" if bgp["doubleprecision"] else "fp32"
    keys = ("g", "eps", "CFL", "theta", "dx", "delta", "lat", "mask", "frictionmodel", "engine", "aoibnd",
            "minlevel", "maxlevel", "Radius", "cf", "dtinit")
    p = default_param(**{k: bgp[k] for k in keys})
    p.update(xo=0.0 if math.isnan(bgp["xo"]) else bgp["xo"], yo=0.0 if math.isnan(bgp["yo"]) else bgp["yo"],
             ForceMassConserve=int(bool(bgp["ForceMassConserve"])), wetdryfix=int(bool(bgp["wetdryfix"])),
             spherical=int(bool(bgp["spherical"])), nblk=int(nblk), nblkmem=int(nblkmem), precision=precision,
             rain_on=int(rain_mmh != 0.0), rain=float(rain_mmh))
    return p
