"""Batches of domains on the GPU: every member of a batch follows exactly the trajectory it has
as a domain of its own (same kernels, its own dt), whatever the other members do, including
members that stop early, members removed and added between runs, and the output path."""
import numpy as np
import pytest

import bg_flood_amd as bgf
from bg_flood_amd import scenarios
from tests._util import domain_for, interior

pytestmark = pytest.mark.gpu


def members(prec, sizes=(64, 96, 48, 128), seed=11):
    out = []
    for k, n in enumerate(sizes):
        sc = scenarios.rain_terrain(n=n, dx=5.0, amp=10.0 + 5 * k, rain=20.0 + 30 * k, manning=0.02 + 0.01 * k,
                                    seed=seed + k, precision="fp64" if prec == "fp64" else "fp32")
        # some water to start with, so the first steps move
        T = np.float64 if prec == "fp64" else np.float32
        m = interior(sc)
        sc["h"] = np.where(m, np.maximum(T(8.0 + k) - sc["zb"], 0), 0).astype(T)
        out.append(sc)
    return out


def solo(sc, prec, endtime, nsteps):
    d = domain_for(sc, precision=prec)
    L = bgf.Loop(nextoutputtime=endtime, dt=0, dtmax=0, totaltime=0)
    d.init_dt(L)
    d.run(L, nsteps, endtime=endtime)
    return d, L


def add_all(B, scs, endtimes):
    return [B.add(sc["param"], sc["blocks"], sc["zb"], sc["h"], sc["u"], sc["v"], sc["cf"], endtime=e)
            for sc, e in zip(scs, endtimes)]


def assert_member_equals(B, m, sc, d, L):
    Lb, done = B.loop(m)
    assert Lb.totaltime == L.totaltime and Lb.dt == L.dt and Lb.dtmax == L.dtmax
    mask = interior(sc)
    a, b = B.get_state(m), d.get_state()
    for k in ("h", "u", "v", "zs"):
        assert np.array_equal(a[k][mask], b[k][mask]), k


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp32_fast"])
def test_batch_members_match_solo_runs(prec):
    scs = members(prec)
    print("case batch %s members=%d" % (prec, len(scs)), flush=True)
    common = dict(scs[0]["param"], precision=prec)
    B = bgf.Batch(common, max_blocks=sum(sc["param"]["nblkmem"] for sc in scs) + 40, max_members=8)
    ids = add_all(B, scs, [1e9] * len(scs))
    B.run(37)
    for m, sc in zip(ids, scs):
        d, L = solo(sc, prec, 1e9, 37)
        assert_member_equals(B, m, sc, d, L)
        assert B.volume(m) == d.volume()
    st = B.stats()
    assert st["cell_updates"] == 37 * sum(sc["param"]["nblk"] for sc in scs) * 256


def test_batch_members_stop_at_their_own_end_times():
    prec = "fp32_fast"
    scs = members(prec)
    ends = [3.0, 1e9, 1.7, 5.5]
    B = bgf.Batch(dict(scs[0]["param"], precision=prec), max_blocks=4000, max_members=6)
    ids = add_all(B, scs, ends)
    B.run(23)
    B.run(18)  # continues the members that are not done
    for m, sc, e in zip(ids, scs, ends):
        d = domain_for(sc, precision=prec)
        L = bgf.Loop(nextoutputtime=e, dt=0, dtmax=0, totaltime=0)
        d.init_dt(L)
        d.run(L, 23, endtime=e)
        d.run(L, 18, endtime=e)
        assert_member_equals(B, m, sc, d, L)
        Lb, done = B.loop(m)
        assert done == (e < 1e9)
        if done:
            assert abs(Lb.totaltime - e) < 1e-6 * e  # the output-time landing ends on endtime


def test_batch_remove_add_and_outputs():
    prec = "fp32"
    scs = members(prec)
    B = bgf.Batch(dict(scs[0]["param"], precision=prec), max_blocks=2200, max_members=4)
    ids = add_all(B, scs[:3], [1e9] * 3)
    B.run(10)
    B.remove(ids[1])
    m3 = B.add(scs[3]["param"], scs[3]["blocks"], scs[3]["zb"], scs[3]["h"], scs[3]["u"], scs[3]["v"],
               scs[3]["cf"], endtime=1e9)
    cells = [(i + 1) + (j + 1) * 18 + b * 324 for i, j, b in ((3, 4, 0), (10, 12, 5), (0, 15, 17))]
    B.set_gauges(m3, cells)
    hm = np.zeros(scs[3]["param"]["nblkmem"] * 324, np.float32)
    for _ in range(3):
        B.run(5)
        B.update_maxdepth()
        hm = np.maximum(hm, B.get_state(m3)["h"].astype(np.float32))
        vals, times = B.sample_gauges()
        st = B.get_state(m3)
        assert np.array_equal(vals[m3, :3], st["h"][cells].astype(np.float32))
        assert np.isnan(vals[m3, 3:]).all()
        assert times[m3] == B.loop(m3)[0].totaltime
    mask = interior(scs[3])
    assert np.array_equal(B.maxdepth(m3)[mask], hm[mask])
    # the survivors and the newcomer match solo runs of the same step counts
    for m, sc, n in ((ids[0], scs[0], 25), (ids[2], scs[2], 25), (m3, scs[3], 15)):
        d, L = solo(sc, prec, 1e9, n)
        assert_member_equals(B, m, sc, d, L)


def test_batch_capacity_errors():
    scs = members("fp32", sizes=(64, 64))
    B = bgf.Batch(dict(scs[0]["param"], precision="fp32"), max_blocks=scs[0]["param"]["nblkmem"] + 3, max_members=4)
    add_all(B, scs[:1], [1e9])
    with pytest.raises(bgf.BgfError):
        add_all(B, scs[1:], [1e9])
