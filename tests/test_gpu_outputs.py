"""A run with map outputs as the reference's main loop makes them (src/Mainloop.cu:60-85): steps
landing on each output time, the statistics kept on the GPU, one NetCDF record per output with the
state and the statistics (Save2Netcdf), the means reset after each output (resetmeanmax)."""
import numpy as np
import pytest
from scipy.io import netcdf_file

import bg_flood_amd as bgf
from bg_flood_amd import ncio, scenarios

pytestmark = pytest.mark.gpu


def test_run_with_map_outputs(tmp_path):
    print("case run with map outputs", flush=True)
    sc = scenarios.rain_terrain(n=128, dx=5.0, amp=4.0, rain=2000.0, manning=0.03, seed=3, precision="fp32")
    p = dict(sc["param"], precision="fp32")
    d = bgf.Domain(p, sc["blocks"])
    zb = np.asarray(sc["zb"], np.float32)
    d.set_state(zb, *[np.asarray(sc[k], np.float32) for k in ("h", "u", "v", "cf")])
    d.set_stats(mean=True, max=True, twet=True, wet_threshold=0.001)
    names = ["zs", "h", "u", "v", "hmax", "zsmax", "hmean", "twet"]
    every, nout = 20.0, 4
    path = tmp_path / "run.nc"
    m = bgf.interior_mask(p["nblkmem"], sc["blocks"]["active"][:p["nblk"]])
    saved = []
    try:
        loop = d.init_dt(bgf.Loop(nextoutputtime=every))
        with ncio.BuqOutput(path, p, sc["blocks"], names) as out:
            for k in range(nout):
                d.run(loop, 100000, endtime=loop.nextoutputtime)  # stops on the output time
                assert loop.totaltime == pytest.approx(every * (k + 1), abs=1e-9)
                st = d.get_state()
                fields = dict(st, **{n: d.stat(n) for n in ("hmax", "zsmax", "hmean", "twet")})
                out.write(loop.totaltime, fields)
                saved.append({n: fields[n].copy() for n in names})
                d.reset_stats(mean=True)
                loop.nextoutputtime += every
    finally:
        d.close()
    with netcdf_file(str(path), "r", mmap=False) as nc:
        t = nc.variables["time"].data.copy()
        grids = {n: nc.variables[n + "_P0"].data.copy() for n in names}
    np.testing.assert_array_equal(t, np.float32(every) * np.arange(1, nout + 1, dtype=np.float32))
    bx = np.asarray(sc["blocks"]["xo"])
    by = np.asarray(sc["blocks"]["yo"])
    ib = sc["blocks"]["active"][5]
    x0, y0 = int(round((bx[ib] - 2.5) / 5.0)), int(round((by[ib] - 2.5) / 5.0))
    for k in range(nout):
        for n in names:
            # the record holds block 5's interior at its place, as the arrays the run gave
            blk = saved[k][n].reshape(-1, 18, 18)[ib, 1:17, 1:17].astype(np.float32)
            np.testing.assert_array_equal(grids[n][k, y0:y0 + 16, x0:x0 + 16], blk, err_msg="%s %d" % (n, k))
    # maxima never decrease, wet time grows with the outputs, means are depths
    assert np.all(np.diff(grids["hmax"], axis=0) >= 0)
    assert np.all(np.diff(grids["twet"], axis=0) >= 0) and grids["twet"][-1].max() > 0
    assert np.all(grids["hmean"][-1] <= grids["hmax"][-1] + 1e-6)
