"""BG_param.txt reading: the reference's line rules (src/ReadInput.cu:1689-1734, 2075-2091) and the
sanity rules for the step's parameters (:1189-1440), on the text of the reference's example
parameter files (Examples/*/BG_param.txt, quoted in part) and on edge cases whose outcome follows
from those rules."""
import math
import os

import pytest

import bg_flood_amd as bgf
from bg_flood_amd import params

CONICAL = """#################################
# Bathymetry file
bathy = Bathy_CI.asc;
posdown=-1;
gpudevice = -1;
eps = 0.00010000;
cf=0.0001;
frictionmodel =0;
CFL = 0.500000;
outputtimestep = 1.000000;
outvars = zs,hh,uu,vv,zb;
endtime = 20.000000;
left =2
TSnode = 4.521, 1.196; #Note these are x, y values converted to the nearest i,j
"""

TPAC = """bathy = Tpac_big.asc;
grdalpha = 0.000000;
spherical = 1;
#xmin=-190.0
ymin=-78.0
ymax=14.32
dx=0.08;
mask = 30.0;
eps = 0.000100;
cf = 0.000100;
theta = 1.300000;
CFL = 0.500000;
totaltime = 0.000000; # Start time
endtime = 54000.000000;
outputtimestep = 600.000000;
g = 9.810000;
"""

SIMPLERAIN = """dem = PortCharles_DEM.nc?z
dx = 32.0;
rain = rain.txt
frictionmodel = 1
roughness = 0.04
engine = 5
endtime = 36000.0
"""


def test_conical_example():
    p = params.read_bg_param(CONICAL)
    assert p["eps"] == 0.0001 and p["cf"] == 0.0001 and p["frictionmodel"] == 0
    assert p["CFL"] == 0.5 and p["outputtimestep"] == 1.0 and p["endtime"] == 20.0
    assert p["engine"] == 1  # the reference's default engine
    assert math.isnan(p["dx"]) and math.isnan(p["delta"])  # dx comes from the bathymetry there
    assert p["minlevel"] == p["maxlevel"] == p["initlevel"] == 0


def test_tpac_example_spherical():
    p = params.read_bg_param(TPAC)
    assert p["spherical"] is True and p["dx"] == 0.08
    assert p["delta"] == 0.08 * 6371220.0 * 3.14159265 / 180.0
    assert p["yo"] == -78.0 and math.isnan(p["xo"])  # '#xmin=' is a comment line: key "#xmin"
    assert p["mask"] == 30.0 and p["theta"] == 1.3 and p["g"] == 9.81
    assert p["totaltime"] == 0.0 and p["inittime"] == 0.0 and p["endtime"] == 54000.0


def test_simplerain_example_engine5_and_roughness_alias():
    p = params.read_bg_param(SIMPLERAIN)
    assert p["engine"] == 5 and p["CFL"] == 0.5
    assert p["cf"] == 0.04 and p["frictionmodel"] == 1 and p["dx"] == 32.0


@pytest.mark.parametrize("line,key,value", [
    ("EPS = 0.01", "eps", 0.01),                 # case-insensitive key
    ("eps==0.02", "eps", 0.02),                  # empty pieces between '=' are dropped
    ("eps = 0.03 = 7", "eps", 0.03),             # later '=' pieces join the value, stod stops at ' '
    ("eps =#0.04", "eps", 0.04),                 # an empty piece before '#' is dropped
    ("eps = #0.05", "eps", 0.0001),              # the piece before '#' is blank: no value
    ("eps = 0.06;0.07", "eps", 0.06),            # cut at ';'
    ("\teps = 0.08", "eps", 0.0001),             # trim is spaces only: key "\teps" is unknown
    ("  eps   = 9e-3 ", "eps", 0.009),
    ("cf = 0.002", "cf", 0.002),
    ("cf = 2e-3", "cf", 0.0001),                 # a letter in the value means a roughness map
    ("cfmap = rough.nc", "cf", 0.0001),
    ("frictionmodel = -1.7", "frictionmodel", -1),  # stoi takes the integer prefix
    ("end = 100", "endtime", 100.0),
    ("stop_time = 12.5", "endtime", 12.5),
    ("start = 3", "totaltime", 3.0),
    ("init = 4", "totaltime", 4.0),
    ("initlevel = 0", "totaltime", 0.0),         # whole-key comparison: not 'init'
    ("outtimestep = 60", "outputtimestep", 60.0),
    ("remainderbnd = 3", "aoibnd", 3),
    ("geo = yes", "spherical", True),
    ("fixinstab = off", "wetdryfix", False),
    ("reminstab = maybe", "wetdryfix", True),    # neither true nor false: unchanged
    ("VolumeCon = On", "ForceMassConserve", True),
    ("engine = b", "engine", 1),
    ("engine = Buttinger", "engine", 1),
    ("engine = kurganov", "engine", 2),
    ("engine = 2", "engine", 2),
    ("engine = 3", "engine", 5),
    ("initzs = 0.5", "zsinit", 0.5),
    ("xmin = 10", "xo", 10.0),
    ("lat = 45", "lat", 0.0),                    # not read from the file
])
def test_line_rules(line, key, value):
    p = params.parse_lines(line)
    assert p[key] == value, (line, p[key])


def test_last_value_wins_and_levels():
    p = params.read_bg_param("maxlevel = 3\nmaxlevel = 2\n")
    assert (p["minlevel"], p["maxlevel"], p["initlevel"]) == (0, 2, 0)
    # an unset minlevel takes initlevel; the two are then ordered
    p = params.read_bg_param("maxlevel = 3\ninitlevel = 7\nmaxlevel = 2\n")
    assert (p["minlevel"], p["maxlevel"], p["initlevel"]) == (2, 7, 7)
    p = params.read_bg_param("minlevel = 3\nmaxlevel = 1\n")
    assert (p["minlevel"], p["maxlevel"], p["initlevel"]) == (1, 3, 1)


def test_sanity_defaults():
    p = params.read_bg_param("dx = 2\n")
    assert p["endtime"] == 1.0 / 0.0000001 and p["zsoffset"] == 0.0 and p["delta"] == 2.0
    assert math.isnan(p["zsinit"])
    p = params.read_bg_param("engine = 5\nCFL = 0.1\n")
    assert p["CFL"] == 0.25
    p = params.read_bg_param("engine = 2\nCFL = 0.1\n")
    assert p["CFL"] == 0.1


def test_bad_values_are_errors():
    with pytest.raises(ValueError):
        params.parse_lines("eps = abc")
    with pytest.raises(ValueError):
        params.parse_lines("endtime = 2020-01-01T00:00:00")


def test_domain_param_builds_the_c_struct(tmp_path):
    f = tmp_path / "BG_param.txt"
    f.write_text(TPAC + "engine = kurganov\ndoubleprecision = 1\n")
    p = params.read_bg_param(path=str(f))
    dp = params.domain_param(p, nblk=4, nblkmem=5)
    assert dp["precision"] == "fp64" and dp["spherical"] == 1 and dp["engine"] == 2
    assert dp["delta"] == p["delta"] and dp["yo"] == -78.0 and dp["xo"] == 0.0
    P = bgf._param_struct(dp)
    assert P.spherical == 1 and P.Radius == 6371220.0 and P.nblk == 4 and P.nblkmem == 5
    with pytest.raises(ValueError):
        params.domain_param(params.read_bg_param(CONICAL), 1, 1)


def test_full_key_set_examples():
    r = params.read_bg_param_all(CONICAL)
    assert r["forcing"]["bathy"] == [dict(file="Bathy_CI.asc", var="", ext="asc")]
    assert r["forcing"]["boundaries"] == [dict(side="left", type=2, file=None)]
    # outvars keeps the names the reference supports: hh, uu, vv are dropped (src/ReadInput.cu:514-522)
    assert r["outvars"] == ["zs", "zb"] and not r["outmax"]
    assert r["param"]["GPUDEVICE"] == -1 and r["param"]["posdown"] == -1
    s = params.read_bg_param_all(SIMPLERAIN)
    assert s["forcing"]["bathy"] == [dict(file="PortCharles_DEM.nc", var="z", ext="nc")]
    assert s["forcing"]["rain"] == dict(file="rain.txt", var="", ext="txt")
    assert s["param"]["cf"] == 0.04 and s["param"]["engine"] == 5
    t = params.read_bg_param_all(TPAC)
    assert t["param"]["spherical"] and t["param"]["yo"] == -78.0 and t["param"]["ymax"] == 14.32


def test_full_key_set_rules():
    text = """outvars = zs, hmax, twet, Umean, bogus;
TSOutput = gauge1.txt, 10.5, 20.25
TSnodesout = g2.txt,1,2
outzone = zone.nc, 0, 100, 0, 50
river = flow.txt, 1.0, 2.0, 3.0, 4.0
right = 3, bnd_right.txt
top = bnd_top.nc,2
cf = roughness.nc?n
il = 5.0
hotstartfile = hot.nc
hotstep = 3
resetmax = 1
smallnc = 1; scalefactor = 0.001
Adaptation = Threshold, zb, 10
savebyblk = no
crs = EPSG:2193
somethingelse = 1
"""
    r = params.read_bg_param_all(text)
    assert r["outvars"] == ["zs", "hmax", "twet", "Umean"]
    assert r["outmean"] and r["outmax"] and r["outtwet"]
    assert r["TSnodesout"] == [dict(name="gauge1.txt", x=10.5, y=20.25), dict(name="g2.txt", x=1.0, y=2.0)]
    assert r["outzone"][0]["name"] == "zone.nc" and r["outzone"][0]["xend"] == 100.0
    assert r["forcing"]["rivers"] == [dict(file="flow.txt", xstart=1.0, xend=2.0, ystart=3.0, yend=4.0)]
    assert dict(side="right", type=3, file="bnd_right.txt") in r["forcing"]["boundaries"]
    assert dict(side="top", type=2, file="bnd_top.nc") in r["forcing"]["boundaries"]
    assert r["forcing"]["cfmap"] == dict(file="roughness.nc", var="n", ext="nc")
    p = r["param"]
    assert p["il"] == 5.0 and p["hotstartfile"] == "hot.nc" and p["hotstep"] == 3 and p["resetmax"]
    assert p["smallnc"] == 1  # the ';' cut keeps the first piece only
    assert p["scalefactor"] == 0.01
    assert p["Adaptation"] == ["Threshold", "zb", "10"]
    assert p["savebyblk"] is False and p["crs_ref"] == "EPSG:2193"
    assert r["unused"] == ["somethingelse = 1"]
    # "file, type" with a space before the type: the reference's isdigit test fails and stoi of the
    # file name throws (src/ReadInput.cu:1960-1972)
    with pytest.raises(ValueError):
        params.read_bg_param_all("top = bnd_top.nc, 2")


def test_timeseries_reader():
    a = params.read_timeseries(text="# rain\n0.0\t317.0\n45.0 317.0\n\n45.1,0.0\n150.0 0.0")
    assert a.shape == (4, 2) and a[2, 0] == 45.1 and a[0, 1] == 317.0


def test_asc_reader_rules():
    body = "1 2 3\n4 5 6\n"
    x, y, z = params.read_asc(text="ncols 3\nnrows 2\nxllcorner 10\nyllcorner 20\ncellsize 2\nnodata_value -9999\n" + body)
    # corner-registered: first node half a cell in; rows stored from the south
    assert list(x) == [11.0, 13.0, 15.0] and list(y) == [21.0, 23.0]
    assert z.tolist() == [[4, 5, 6], [1, 2, 3]]
    x, y, z = params.read_asc(text="ncols 3\nnrows 2\nxllcenter 10\nyllcenter 20\ncellsize 2\nnodata_value -9999\n" + body)
    # the reference's rule: yllcenter also marks the grid corner-registered
    assert list(x) == [11.0, 13.0, 15.0] and list(y) == [21.0, 23.0]


RAIN_CEA = "/root/reference/Examples/Rain_Cea2008"


@pytest.mark.skipif(not os.path.exists(RAIN_CEA), reason="the reference's example data is not present")
def test_reference_rain_example_files():
    r = params.read_bg_param_all(path=os.path.join(RAIN_CEA, "BG_param.txt"))
    assert r["outvars"] == ["zs", "zb", "zsmax"] and r["outmax"]
    assert r["forcing"]["boundaries"] == [dict(side="bot", type=1, file=None)]
    assert r["param"]["doubleprecision"] == 1 and r["param"]["theta"] == 1.14 and r["param"]["frictionmodel"] == 1
    s = params.read_timeseries(os.path.join(RAIN_CEA, r["forcing"]["rain"]["file"]))
    assert s.tolist() == [[0.0, 317.0], [45.0, 317.0], [45.1, 0.0], [150.0, 0.0]]
    x, y, z = params.read_asc(os.path.join(RAIN_CEA, "topo.asc"))
    assert z.shape == (256, 208) and abs(x[1] - x[0] - 0.01) < 1e-12
