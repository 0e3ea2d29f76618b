"""In-run adaptation on the GPU: flags from the GPU state, the adaptation pass with coarsening, and
the state moved onto the new mesh by bgf_remesh, against the oracle's restatement, bit for bit, over
several adapt-and-step cycles; with the bed of new cells from the state or from a bed raster."""
import numpy as np
import pytest

import bg_flood_amd as bgf
from bg_flood_amd import scenarios
from oracle import pyoracle as po
from tests._util import domain_for, interior, oracle_for, orc_param_with

pytestmark = pytest.mark.gpu

NBX, NBY, DX = 6, 6, 8.0


def bed(x, y):
    return -6.0 + 0.004 * x + 8.0 * np.exp(-((x - 300.0) ** 2 + (y - 250.0) ** 2) / 80.0 ** 2)


def scenario(prec, maxlevel=3):
    T = np.float64 if prec == "fp64" else np.float32
    blocks, nblk, nblkmem = scenarios.adapt_mesh(NBX, NBY, DX, 2, lambda x, y: np.hypot(x - 380, y - 380) < 70,
                                                 precision=prec)
    p = bgf.default_param(dx=DX, delta=DX, nblk=nblk, nblkmem=nblkmem, minlevel=0, maxlevel=maxlevel,
                          precision=prec, aoibnd=1, frictionmodel=0, cf=0.002)
    X, Y = scenarios.cell_coords_levels(blocks, nblkmem, 0.0, 0.0, DX, np.float64)
    zb = bed(X, Y)
    zs = 0.6 * np.exp(-((X - 380.0) ** 2 + (Y - 380.0) ** 2) / 40.0 ** 2)
    h = np.maximum(zs - zb, 0.0)
    sc = dict(param=p, blocks=blocks, nbx=NBX, nby=NBY)
    for k, a in (("zb", zb), ("h", h), ("u", 0 * h), ("v", 0 * h), ("cf", 0 * h + 0.002)):
        sc[k] = a.reshape(-1).astype(T)
    return sc


def raster():
    g = np.arange(0.0, NBX * 16 * DX + 1e-9, 4.0)
    Xg, Yg = np.meshgrid(g, g)
    return bed(Xg, Yg).astype(np.float32), 0.0, 0.0, 4.0


def steps(d, o, lg, lo, n):
    for k in range(n):
        d.step(lg)
        o.step(lo)
        assert lg.dt == lo.dt, "step %d" % k
        lg.totaltime += lg.dt
        lo.totaltime += lo.dt


def assert_same_state(d, o):
    m = bgf.interior_mask(d.nblkmem, d.active)
    st = d.get_state()
    for k in ("h", "u", "v", "zs"):
        a, b = st[k][m], o.get(k)[m]
        assert np.array_equal(a, b), "%s: max diff %g" % (k, np.abs(a - b).max())


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("with_raster", [False, True])
def test_adapt_cycles_match_the_oracle(prec, with_raster):
    sc = scenario(prec)
    d = domain_for(sc, precision=prec)
    o = oracle_for(sc, precision=prec)
    rs = raster() if with_raster else None
    if rs:
        d.set_bed_source(rs[0], rs[1], rs[2], rs[3])
    lg = bgf.Loop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    lo = po.OrcLoop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    d.init_dt(lg)
    o.init_loop(lo)
    crit = [("zs", 2, 2e-3, 0.0), ("h", 1, 1e-4, 0.05)]
    history = []
    for cycle in range(4):
        print("case remesh %s raster=%s cycle=%d nblk=%d" % (prec, with_raster, cycle, d.nblk), flush=True)
        steps(d, o, lg, lo, 10)
        assert_same_state(d, o)
        before = d.nblk
        ref = np.zeros(d.nblkmem, np.uint8)
        for f, mode, a, b in crit:
            ref |= d.refine_flags(f, mode, a, b)
        co = np.zeros(d.nblkmem, np.uint8)
        co[d.active] = 1 - ref[d.active]
        blocks, nb, nm, remap = bgf.adapt_mesh(d.param, d.blocks, ref, co, NBX, NBY)
        d.remesh(blocks, nb, nm, remap)
        o.remesh(orc_param_with(o.p, nblk=nb, nblkmem=nm), blocks, remap, bed=rs)
        kinds = np.bincount(remap[blocks["active"][:nb], 0], minlength=3)
        history.append((before, nb, tuple(kinds)))
        assert_same_state(d, o)
    steps(d, o, lg, lo, 10)
    assert_same_state(d, o)
    print("history", history, flush=True)
    # the mesh did change both ways
    assert any(k[1] > 0 for _, _, k in history) and any(k[2] > 0 for _, _, k in history)


def test_adapt_helper_keeps_the_lake_at_rest_away_from_the_shore():
    """Domain.adapt (flags + pass + remesh in one call) on still water over the bed raster: the
    interpolated surface is the still level wherever the interpolation stencil is all wet (here:
    every cell deeper than 1 m), velocities stay zero; at the shoreline the reference's
    interpolation of zs mixes wet and dry cells (src/Adaptation.cu:1032-1102)."""
    sc = scenario("fp64")
    sc["h"] = np.maximum(0.0 - sc["zb"], 0.0)
    d = domain_for(sc, precision="fp64")
    d.set_bed_source(*raster())
    n0, n1, nref, nco = d.adapt([("zb", 1, -0.5, 0.5)], NBX, NBY, coarsen=True)
    print("case adapt helper %d -> %d (%d refined, %d coarsened)" % (n0, n1, nref, nco), flush=True)
    assert nref > 0 and n1 > n0
    st = d.get_state()
    m = bgf.interior_mask(d.nblkmem, d.active)
    zb = st["zs"][m] - st["h"][m]
    deep = zb < -1.0
    assert deep.sum() > 1000
    assert np.max(np.abs(st["zs"][m][deep])) < 1e-12
    assert np.all(st["u"][m] == 0) and np.all(st["v"][m] == 0)
