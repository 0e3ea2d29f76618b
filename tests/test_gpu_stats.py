"""Output statistics on the GPU (Calcmeanmax, src/Meanmax.cu:15-112: means, maxima, wet time)
against the oracle's restatement, bit for bit with the EXACT kernels, over steps that cross output
times (the means are divided at the step that reaches nextoutputtime, then reset as resetmeanmax
does), through single steps (the caller advances totaltime) and through bgf_run (the GPU does)."""
import numpy as np
import pytest

import bg_flood_amd as bgf
from bg_flood_amd import scenarios
from oracle import pyoracle as po
from tests._util import domain_for, interior, oracle_for

pytestmark = pytest.mark.gpu

NAMES = po.Stats.NAMES


def wet_dry_case(prec):
    # a hump of water running over a bed with a dry island: wet/dry cells for the wet time,
    # velocities for U, hU, hUU
    sc = scenarios.gaussian_hump(n=64, dx=10.0, depth=1.0, amp=1.5, radius=80.0, precision=prec, aoibnd=1)
    X, Y = scenarios.cell_coords(sc["blocks"], sc["param"]["nblkmem"], 0.0, 0.0, 10.0, np.float64)
    T = np.float64 if prec == "fp64" else np.float32
    zb = (-1.0 + 1.6 * np.exp(-((X - 480.0) ** 2 + (Y - 200.0) ** 2) / 90.0 ** 2)).astype(T)
    zs = np.maximum(zb, (0.0 + 1.5 * np.exp(-((X - 320.0) ** 2 + (Y - 320.0) ** 2) / (2 * 80.0 ** 2))).astype(T))
    h = np.maximum(zs - zb, 0).astype(T)
    m = interior(sc)
    h[~m] = 0
    sc.update(zb=zb, h=h)
    return sc


def compare(d, st, m, tag):
    for name in NAMES:
        try:
            a = d.stat(name)
        except bgf.BgfError:
            continue
        b = st.get(name)
        bad = np.flatnonzero(a[m] != b[m])
        assert bad.size == 0, (tag, name, bad[:5], a[m][bad[:5]], b[m][bad[:5]])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_stats_single_steps(prec):
    print("case stats single steps", prec, flush=True)
    sc = wet_dry_case(prec)
    o = oracle_for(sc, precision=prec)
    d = domain_for(sc, precision=prec)
    m = interior(sc)
    st = po.Stats(o, 7, wet_threshold=0.1)
    d.set_stats(mean=True, max=True, twet=True, wet_threshold=0.1)
    out_every = 2.0
    lo = po.OrcLoop(nextoutputtime=out_every, dt=0, dtmax=0, totaltime=0)
    lg = bgf.Loop(nextoutputtime=out_every, dt=0, dtmax=0, totaltime=0)
    o.init_loop(lo)
    d.init_dt(lg)
    outputs = 0
    wet_and_dry = False
    for k in range(40):
        o.step(lo)
        d.step(lg)
        assert lg.dt == lo.dt, (k, lg.dt, lo.dt)
        lo.totaltime += lo.dt
        lg.totaltime += lg.dt
        st.update(lo)
        if lo.nextoutputtime - lo.totaltime <= lo.dt * 1e-5 or k % 9 == 8:
            compare(d, st, m, (prec, k))
            tw = d.stat("twet")[m]
            wet_and_dry = wet_and_dry or (np.any(tw > 0) and np.any(tw == 0))
        if lo.nextoutputtime - lo.totaltime <= lo.dt * 1e-5:
            # output produced: reset the means (and with resetmax the maxima and wet time)
            outputs += 1
            which = 1 | (6 if outputs % 2 == 0 else 0)
            st.reset(which)
            d.reset_stats(mean=True, max=bool(which & 2), twet=bool(which & 4))
            lo.nextoutputtime += out_every
            lg.nextoutputtime += out_every
    assert outputs >= 2
    compare(d, st, m, (prec, "end"))
    assert wet_and_dry  # the wet time saw both wet and dry cells


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_stats_through_run(prec):
    """bgf_run (graph replays, totaltime advanced on the GPU) gives the statistics of the same steps
    taken one by one in the oracle, the division at the output step included."""
    print("case stats run", prec, flush=True)
    sc = wet_dry_case(prec)
    o = oracle_for(sc, precision=prec)
    d = domain_for(sc, precision=prec)
    m = interior(sc)
    st = po.Stats(o, 3)
    d.set_stats(mean=True, max=True)
    lo = po.OrcLoop(nextoutputtime=3.0, dt=0, dtmax=0, totaltime=0)
    lg = bgf.Loop(nextoutputtime=3.0, dt=0, dtmax=0, totaltime=0)
    o.init_loop(lo)
    d.init_dt(lg)
    n = 0
    while lo.nextoutputtime - lo.totaltime > lo.dt * 1e-5:
        o.step(lo)
        lo.totaltime += lo.dt
        st.update(lo)
        n += 1
    d.run(lg, n)
    assert lg.totaltime == lo.totaltime
    compare(d, st, m, (prec, "run"))
    # the means were divided at the output step: the mean depth is a depth
    assert np.abs(d.stat("hmean")[m] - o.get("h")[m]).max() < 1.0


def test_stats_off_and_errors():
    sc = wet_dry_case("fp32")
    d = domain_for(sc, precision="fp32")
    with pytest.raises(bgf.BgfError):
        d.stat("hmax")
    d.set_stats(max=True)
    with pytest.raises(bgf.BgfError):
        d.stat("hmean")
    with pytest.raises(bgf.BgfError):
        d.stat("nothing")
    assert np.all(d.stat("hmax")[interior(sc)] == -np.finfo(np.float32).max)
