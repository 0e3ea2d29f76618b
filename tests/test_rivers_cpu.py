"""River discharges on the oracle (AddRiverForcing, InjectRiverCPU, src/Updateforcing.cu:146-363;
discharge cells and area, src/InitialConditions.cu:346-445), pinned by closed forms: a still, walled
basin gains exactly the interpolated flow times the time step, spread over the discharge cells."""
import numpy as np
import pytest

from bg_flood_amd import scenarios
from oracle import pyoracle as po
from tests._util import oracle_for, interior


def basin(prec, n=64, dx=4.0):
    sc = scenarios.gaussian_hump(n=n, dx=dx, depth=2.0, amp=0.0, precision=prec)
    return sc


def run(o, nsteps):
    L = po.OrcLoop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    o.init_loop(L)
    rec = []
    for _ in range(nsteps):
        t0 = L.totaltime
        o.step(L)
        rec.append((t0, L.dt))
        L.totaltime += L.dt
    return rec


def test_volume_gained_is_the_flow():
    sc = basin("fp64")
    o = oracle_for(sc)
    times, flows = [0.0, 5.0, 100.0], [0.0, 8.0, 2.0]
    # a rectangle over 3 x 2 cells well inside the domain (cell centres at 2 + 4k)
    o.set_rivers([(99.0, 109.0, 131.0, 135.0, times, flows)])
    v0 = o.volume()
    rec = run(o, 50)
    want = v0 + sum(np.interp(t0, times, flows) * dt for t0, dt in rec)
    assert abs(o.volume() - want) <= 1e-12 * v0
    # the discharge cells: columns with centre 98..110 -> boxes [96,100],[100,104],... overlap [99,109]
    h = scenarios.grid_from_buq(o.get("h"), 4, 4)
    assert h.max() > 2.0 and rec[-1][0] > 5.0


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_one_step_adds_the_flow_at_the_discharge_cells(prec):
    """One step from still water: the volume gained is Q*dt, and the two discharge cells (boxes
    [40, 44] x [40, 44] and [44, 48] x [40, 44] overlap the rectangle) gain the most."""
    n, dx = 32, 4.0
    sc = basin(prec, n=n, dx=dx)
    o = oracle_for(sc, precision=prec)
    Q = 3.0
    o.set_rivers([(40.5, 47.5, 40.5, 43.5, [0.0, 1e6], [Q, Q])])  # cells 10, 11 in x; 10 in y: 2 cells
    (t0, dt), = run(o, 1)
    h = scenarios.grid_from_buq(o.get("h"), 2, 2)
    T = np.float64 if prec == "fp64" else np.float32
    gained = (h - T(2.0)).astype(np.float64)
    top2 = sorted(zip(gained.ravel(), range(gained.size)), reverse=True)[:2]
    assert sorted(k for _, k in top2) == [10 * n + 10, 10 * n + 11]
    assert abs(np.sum(gained) * dx * dx - Q * dt) < (1e-12 if prec == "fp64" else 1e-5) * Q * dt


def test_river_outside_the_domain_is_dropped():
    sc = basin("fp64")
    a = oracle_for(sc)
    b = oracle_for(sc)
    b.set_rivers([(-50.0, -40.0, 10.0, 20.0, [0.0, 1e6], [5.0, 5.0])])
    run(a, 10)
    run(b, 10)
    assert np.array_equal(a.get("h"), b.get("h"))
