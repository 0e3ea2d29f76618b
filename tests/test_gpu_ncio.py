"""Gridded inputs and hot starts on the GPU: bgf_sample_raster (interp2BUQ onto the mesh) against
the oracle's restatement bit for bit, on a multi-level mesh with rasters finer and coarser than its
cells; and a hot start (readhotstartfile, src/InitEvolv.cu:598-835) from the library's own BUQ map
output, which reproduces the state that was written."""
import numpy as np
import pytest

import bg_flood_amd as bgf
from bg_flood_amd import ncio, scenarios
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

NBX, NBY, DX = 6, 6, 8.0


def multilevel(prec):
    blocks, nblk, nblkmem = scenarios.adapt_mesh(NBX, NBY, DX, 2, lambda x, y: np.hypot(x - 380, y - 380) < 70,
                                                 precision=prec)
    p = bgf.default_param(dx=DX, delta=DX, xo=5.0, yo=-3.0, nblk=nblk, nblkmem=nblkmem, minlevel=0, maxlevel=2,
                          precision=prec, aoibnd=1)
    return p, blocks


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("rdx", [3.0, 11.0, 1.5])
def test_sample_raster_bitexact(prec, rdx):
    print("case sample_raster", prec, rdx, flush=True)
    p, blocks = multilevel(prec)
    rng = np.random.default_rng(int(rdx * 10))
    n = int(NBX * 16 * DX / rdx) + 4
    vals = (rng.standard_normal((n, n)) * 5).astype(np.float32)
    xo, yo = 0.0, -7.0
    d = bgf.Domain(p, blocks)
    try:
        got = d.sample_raster(vals, xo, yo, rdx)
    finally:
        d.close()
    op = po.make_param(dx=DX, delta=DX, xo=5.0, yo=-3.0, nblk=p["nblk"], nblkmem=p["nblkmem"], minlevel=0, maxlevel=2,
                       doubleprecision=int(prec == "fp64"))
    ref = po.sample_raster(op, blocks, vals, xo, yo, rdx)
    m = bgf.interior_mask(p["nblkmem"], blocks["active"][:p["nblk"]])
    np.testing.assert_array_equal(got[m], ref[m])
    assert np.all(got[~m] == 0)


def run_scenario(prec):
    sc = scenarios.rain_terrain(n=128, precision="fp32")
    p = dict(sc["param"], precision=prec)
    T = np.float64 if prec == "fp64" else np.float32
    d = bgf.Domain(p, sc["blocks"])
    zb = np.asarray(sc["zb"], T)
    d.set_state(zb, *[np.asarray(sc[k], T) for k in ("h", "u", "v", "cf")])
    loop = d.init_dt(bgf.Loop(nextoutputtime=1e30))
    d.run(loop, 60)
    return sc, p, d, zb, loop


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_hotstart_roundtrip(tmp_path, prec):
    print("case hotstart", prec, flush=True)
    sc, p, d, zb, loop = run_scenario(prec)
    try:
        st = d.get_state()
        path = tmp_path / "out.nc"
        with ncio.BuqOutput(path, p, sc["blocks"], ["zb", "zs", "h", "u", "v"]) as o:
            o.write(loop.totaltime, dict(st, zb=zb))
        assert ncio.read_var(path, "time")[0] == np.float32(loop.totaltime)
        hs = ncio.hotstart(d, path, step=0)
    finally:
        d.close()
    m = bgf.interior_mask(p["nblkmem"], sc["blocks"]["active"][:p["nblk"]])
    # the file holds the state as float; sampled at the grid points it is read back unchanged
    for k, ref in (("zb", zb), ("h", st["h"]), ("u", st["u"]), ("v", st["v"])):
        np.testing.assert_array_equal(hs[k][m], ref[m].astype(np.float32).astype(hs[k].dtype), err_msg=k)
    # a hot-started domain continues from the written state
    d2 = bgf.Domain(p, sc["blocks"])
    try:
        d2.set_state(hs["zb"], hs["h"], hs["u"], hs["v"], np.asarray(sc["cf"], hs["h"].dtype))
        loop2 = d2.init_dt(bgf.Loop(nextoutputtime=1e30, totaltime=loop.totaltime))
        d2.run(loop2, 5)
        assert np.isfinite(d2.get_state()["h"][m]).all()
    finally:
        d2.close()


def test_hotstart_zs_only(tmp_path):
    """A file with zb and zs only: h = max(zs - zb, 0), u = v = 0 (src/InitEvolv.cu:700-760)."""
    print("case hotstart zs only", flush=True)
    sc, p, d, zb, loop = run_scenario("fp32")
    try:
        st = d.get_state()
        path = tmp_path / "zs.nc"
        with ncio.BuqOutput(path, p, sc["blocks"], ["zb", "zs"]) as o:
            o.write(loop.totaltime, dict(st, zb=zb))
        hs = ncio.hotstart(d, path)
        path2 = tmp_path / "u.nc"
        with ncio.BuqOutput(path2, p, sc["blocks"], ["u"]) as o:
            o.write(loop.totaltime, st)
        assert ncio.hotstart(d, path2) is None  # neither zs nor h: cold start
    finally:
        d.close()
    m = bgf.interior_mask(p["nblkmem"], sc["blocks"]["active"][:p["nblk"]])
    zs32 = st["zs"].astype(np.float32)
    np.testing.assert_array_equal(hs["h"][m], np.maximum(np.maximum(zs32, zb) - zb, np.float32(0))[m])
    assert np.all(hs["u"] == 0) and np.all(hs["v"] == 0)
