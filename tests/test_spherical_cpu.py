"""Spherical grids on the CPU: pins of the oracle's metric terms (src/Spherical.cu) by closed
forms and by invariants of the discrete scheme on the sphere."""
import math

import numpy as np
import pytest

from bg_flood_amd import scenarios
from oracle import pyoracle as po
from tests._util import oracle_for, interior

PI = 3.14159265  # src/General.h:6


@pytest.mark.parametrize("yo,delta", [(0.0, 1000.0), (-35.5, 250.0), (61.2, 4000.0)])
def test_calccm_calcfm_closed_forms(yo, delta):
    """cm = (sin(phi+dphi) - sin(phi-dphi)) / (2 dphi) = cos(phi) sin(dphi)/dphi at the row's upper face;
    fm = cos of the latitude of face row iy."""
    R = 6371220.0
    L = po.lib()
    for iy in range(17):
        phi = (yo + (iy + 0.5) * delta / R * (180.0 / PI)) * (PI / 180.0)
        dphi = delta / (2.0 * R)
        assert L.orc_calccm_d(R, delta, yo, iy) == pytest.approx(math.cos(phi) * math.sin(dphi) / dphi, rel=1e-9)
        lat = yo + iy * (delta / R * (180.0 / PI))
        assert L.orc_calcfm_d(R, delta, yo, float(iy)) == pytest.approx(math.cos(lat * PI / 180.0), rel=1e-13)


def _run(sc, n):
    o = oracle_for(sc)
    L = po.OrcLoop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    o.init_loop(L)
    for _ in range(n):
        o.step(L)
        L.totaltime += L.dt
    return o


def _cm_cells(sc):
    """cm of every interior cell, from the oracle's calcCM (row iy of its block)."""
    p, b = sc["param"], sc["blocks"]
    cm = np.zeros(p["nblkmem"] * 324)
    L = po.lib()
    for ib in range(p["nblk"]):
        lev = b["level"][ib]
        d = p["delta"] / (1 << lev)
        ybo = p["yo"] + b["yo"][ib]
        for iy in range(16):
            c = L.orc_calccm_d(p["Radius"], d, ybo, iy)
            for ix in range(16):
                cm[(ix + 1) + (iy + 1) * 18 + ib * 324] = c * d * d
    return cm


def test_spherical_lake_at_rest():
    """Still water over a flat bed stays at rest: the fm-weighted face pressure fluxes are balanced
    by the metric source term h*(g/2*h*dmdt) (a wrong sign or a missing term moves water).  Neumann
    edges: the reference's wall patch (src/Boundary.cu:1887-1891) zeroes the normal momentum flux
    difference of the edge cell, which on the sphere leaves that cell's metric term unbalanced."""
    sc = scenarios.spherical_basin(amp=0.0, lat0=20.0, dx=0.05, aoibnd=1)
    o = _run(sc, 40)
    m = interior(sc)
    assert np.abs(o.get("u")[m]).max() < 1e-11 and np.abs(o.get("v")[m]).max() < 1e-11
    assert np.abs(o.get("h")[m] - 10.0).max() < 1e-11


@pytest.mark.parametrize("maxlevel", [0, 2])
def test_spherical_volume_is_conserved(maxlevel):
    """With walls the area-weighted volume sum(h * cm * delta^2) changes only by round-off."""
    sc = scenarios.spherical_basin(lat0=30.0, lat_c=30.3, dx=0.01, maxlevel=maxlevel)
    area = _cm_cells(sc)
    o0 = oracle_for(sc)
    v0 = float((o0.get("h") * area).sum())
    o = _run(sc, 40)
    v1 = float((o.get("h") * area).sum())
    assert abs(v1 - v0) / v0 < 1e-12


def test_spherical_first_step_dt_closed_form():
    """Still water: every face has a = sqrt(g h) and the CFL step CFL*delta*cm/fm/a; fm <= 1 on the
    Y faces, so the step is CFL*delta*min(cm)/sqrt(g h) over the rows of the domain (a missing or
    misplaced cm on the X faces changes it).  (Coriolis and the half-row offset of cm make the
    solution itself neither east-west nor equator symmetric.)"""
    sc = scenarios.spherical_basin(amp=0.0, lat0=40.0, dx=0.05, aoibnd=1)
    p, b = sc["param"], sc["blocks"]
    o = oracle_for(sc)
    L = po.OrcLoop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    o.init_loop(L)
    o.step(L)
    lib = po.lib()
    cms = [lib.orc_calccm_d(p["Radius"], p["delta"], p["yo"] + b["yo"][ib], iy)
           for ib in range(p["nblk"]) for iy in range(16)]
    expect = p["CFL"] * p["delta"] * min(cms) / math.sqrt(p["g"] * 10.0)
    rem = 1e9  # landing on nextoutputtime (src/Advection.cu:525-532)
    expect = rem / math.ceil(rem / expect)
    assert L.dt == pytest.approx(expect, rel=1e-12)
