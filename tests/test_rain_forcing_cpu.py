"""Rain forcing beyond a constant rate, on the oracle: a uniform time series (Forcingthisstep +
interptime, src/Updateforcing.cu:45-60) and rain maps with frames (InterpstepCPU, interp2BUQ,
src/GridManip.cu:520-622; AddrainforcingImplicitCPU, src/Updateforcing.cu:530-581), pinned by
closed forms: a flat, still, walled basin gains exactly the rain each step applies."""
import numpy as np
import pytest

from bg_flood_amd import scenarios
from oracle import pyoracle as po
from tests._util import oracle_for, interior


def basin(prec, n=64, dx=4.0, rain=0.0):
    """Still water 2 m deep over a flat bed, walls, no friction: every cell keeps the same depth and
    the volume changes by the rain alone."""
    sc = scenarios.gaussian_hump(n=n, dx=dx, depth=2.0, amp=0.0, precision=prec)
    sc["param"] = dict(sc["param"], rain_on=1, rain=rain)
    return sc


def run_steps(o, n):
    L = po.OrcLoop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    o.init_loop(L)
    rec = []
    for _ in range(n):
        t0 = L.totaltime
        o.step(L)
        rec.append((t0, L.dt))
        L.totaltime += L.dt
    return rec


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_constant_series_equals_constant_rain(prec):
    a = oracle_for(basin(prec, rain=36.0))
    b = oracle_for(basin(prec, rain=0.0))
    b.set_rain_series([0.0, 1e6], [36.0, 36.0])
    run_steps(a, 12)
    run_steps(b, 12)
    for k in ("h", "u", "v"):
        assert np.array_equal(a.get(k), b.get(k)), k


def test_series_ramp_adds_the_interpolated_rain():
    sc = basin("fp64")
    o = oracle_for(sc)
    times, rates = [0.0, 10.0, 30.0], [0.0, 360.0, 72.0]
    o.set_rain_series(times, rates)
    rec = run_steps(o, 40)
    m = interior(sc)
    h = o.get("h")[m]
    assert np.ptp(h) == 0.0  # still water stays flat
    depth = 2.0
    for t0, dt in rec:
        rate = np.interp(t0, times, rates)  # inside the series the reference's interptime is linear
        depth += rate / 1000.0 / 3600.0 * dt
    assert abs(h[0] - depth) < 1e-13
    assert rec[-1][0] + rec[-1][1] > 10.0  # the ramp's second segment was reached


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_map_at_cell_centres_gives_each_cell_its_rate(prec):
    """A rain map whose nodes are the cell centres: each cell receives exactly its node's rate
    (bilinear weights 1 and 0), so a uniform map equals the constant rate bit for bit."""
    n, dx = 64, 4.0
    T = np.float64 if prec == "fp64" else np.float32
    a = oracle_for(basin(prec, n=n, dx=dx, rain=36.0))
    b = oracle_for(basin(prec, n=n, dx=dx, rain=0.0))
    b.set_rain_map(np.full((n, n), 36.0, np.float32), dx / 2, dx / 2, dx)
    run_steps(a, 6)
    run_steps(b, 6)
    for k in ("h", "u", "v"):
        assert np.array_equal(a.get(k), b.get(k)), k
    # a varying map: after one step, the depth increment of each cell is its node's rate * dt
    rng = np.random.default_rng(4)
    rates = (rng.random((n, n)) * 100).astype(np.float32)
    sc = basin(prec, n=n, dx=dx)
    c = oracle_for(sc)
    c.set_rain_map(rates, dx / 2, dx / 2, dx)
    (t0, dt), = run_steps(c, 1)
    got = scenarios.grid_from_buq(c.get("h"), n // 16, n // 16) - T(2.0)
    want = np.maximum(rates.astype(T) / T(1000.0) / T(3600.0) * T(dt), 0)
    # the first step's fluxes are zero (still water) so only the rain changes the depth
    assert np.max(np.abs(got - want)) <= (1e-15 if prec == "fp64" else 4e-7 * 2.0)


def test_map_frames_interpolate_in_time():
    n, dx = 32, 4.0
    sc = basin("fp64", n=n, dx=dx)
    o = oracle_for(sc)
    f0 = np.full((n, n), 10.0, np.float32)
    f1 = np.full((n, n), 130.0, np.float32)
    f2 = np.full((n, n), 70.0, np.float32)
    o.set_rain_map(np.stack([f0, f1, f2]), dx / 2, dx / 2, dx, t0=0.0, dt=4.0)
    rec = run_steps(o, 60)
    h = o.get("h")[interior(sc)]
    assert rec[-1][0] > 13.0
    # between frames: linear; beyond the last frame the last pair keeps being extrapolated (the frame
    # index is clamped to nframes - 2) down to a negative rate, which the step clamps to no rain
    depth = 2.0
    fr = [10.0, 130.0, 70.0]
    for t0, dt in rec:
        k = min(max(int(np.floor(t0 / 4.0)), 0), 1)
        r = float(np.float32(fr[k] + (t0 - 4.0 * k) * (fr[k + 1] - fr[k]) / 4.0))  # the map is stored in float
        depth += max(r / 1000.0 / 3600.0 * dt, 0.0)
    assert h[0] == depth
