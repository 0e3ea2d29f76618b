"""Rain forcing on the GPU against the oracle, bit for bit: a uniform rain series and rain maps with
frames, on one-level meshes (fused kernels, with the steps kept on the GPU and replayed as a graph)
and on multi-level meshes (general kernels), in FP64 and FP32."""
import numpy as np
import pytest

import bg_flood_amd as bgf
from bg_flood_amd import scenarios
from oracle import pyoracle as po
from tests._util import domain_for, interior, oracle_for

pytestmark = pytest.mark.gpu

SERIES = ([0.0, 5.0, 20.0, 60.0], [0.0, 250.0, 40.0, 120.0])


def frames(n_cells_x, n_cells_y, dx, seed=3):
    rng = np.random.default_rng(seed)
    rdx = 3.0 * dx  # a map coarser than the cells, not aligned with them
    nx, ny = int(n_cells_x * dx / rdx) + 2, int(n_cells_y * dx / rdx) + 2
    f = (rng.random((3, ny, nx)) * 200.0).astype(np.float32)
    return f, -0.3 * dx, -0.2 * dx, rdx


def setup(sc, prec, kind):
    d = domain_for(sc, precision=prec)
    o = oracle_for(sc, precision=prec)
    if kind == "series":
        d.set_rain_series(*SERIES)
        o.set_rain_series(*SERIES)
    else:
        f, xo, yo, rdx = frames(sc["nbx"] * 16, sc["nby"] * 16, sc["param"]["dx"])
        d.set_rain_map(f, xo, yo, rdx, t0=0.0, dt=3.0)
        o.set_rain_map(f, xo, yo, rdx, t0=0.0, dt=3.0)
    return d, o


def compare(sc, d, o):
    m = bgf.interior_mask(d.nblkmem, d.active)
    st = d.get_state()
    for k in ("h", "u", "v"):
        a, b = st[k][m], o.get(k)[m]
        assert np.array_equal(a, b), "%s: max diff %g" % (k, np.abs(a - b).max())


def run_pair(sc, prec, kind, nsteps, graph):
    print("case rain %s %s nblk=%d graph=%s" % (kind, prec, sc["param"]["nblk"], graph), flush=True)
    d, o = setup(sc, prec, kind)
    lg = bgf.Loop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    lo = po.OrcLoop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    d.init_dt(lg)
    o.init_loop(lo)
    if graph:
        d.run(lg, nsteps)
        for _ in range(nsteps):
            o.step(lo)
            lo.totaltime += lo.dt
        assert lg.totaltime == lo.totaltime
    else:
        for k in range(nsteps):
            d.step(lg)
            o.step(lo)
            assert lg.dt == lo.dt, "step %d" % k
            lg.totaltime += lg.dt
            lo.totaltime += lo.dt
    compare(sc, d, o)
    return lo.totaltime


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["series", "map"])
def test_one_level_rain_forcing_exact(prec, kind):
    sc = scenarios.rain_terrain(n=96, dx=5.0, amp=6.0, rain=0.0, manning=0.01, seed=2,
                                precision="fp64" if prec == "fp64" else "fp32")
    # quadratic friction: IEEE operations only (Manning's cube root is not correctly rounded)
    sc["param"] = dict(sc["param"], rain_on=0, frictionmodel=0)
    t = run_pair(sc, prec, kind, 30, graph=False)
    assert t > 5.0
    run_pair(sc, prec, kind, 30, graph=True)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["series", "map"])
def test_multi_level_rain_forcing_exact(prec, kind):
    sc = scenarios.amr_basin(precision=prec, island=True, amp=1.0, frictionmodel=0, cf=0.002)
    run_pair(sc, prec, kind, 25, graph=False)
    run_pair(sc, prec, kind, 25, graph=True)


def test_rain_map_survives_a_remesh():
    from tests._util import orc_param_with
    from tests.test_gpu_remesh import scenario, NBX, NBY
    sc = scenario("fp64")
    d, o = setup(sc, "fp64", "map")
    lg = bgf.Loop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    lo = po.OrcLoop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    d.init_dt(lg)
    o.init_loop(lo)
    for cycle in range(3):
        for _ in range(8):
            d.step(lg)
            o.step(lo)
            assert lg.dt == lo.dt
            lg.totaltime += lg.dt
            lo.totaltime += lo.dt
        ref = d.refine_flags("zs", 2, 2e-3)
        co = np.zeros(d.nblkmem, np.uint8)
        co[d.active] = 1 - ref[d.active]
        blocks, nb, nm, remap = bgf.adapt_mesh(d.param, d.blocks, ref, co, NBX, NBY)
        d.remesh(blocks, nb, nm, remap)
        o.remesh(orc_param_with(o.p, nblk=nb, nblkmem=nm), blocks, remap)
        compare(sc, d, o)
