"""Launch schedules of the persistent step: every tile shape and workgroup count gives the same
results, bit for bit, on domains and batches, and the autotuner returns a schedule that applies."""
import numpy as np
import pytest

import bg_flood_amd as bgf
from bg_flood_amd import autotune, scenarios
from tests._util import domain_for, interior
from tests.test_gpu_batch import add_all, assert_member_equals, members, solo

pytestmark = pytest.mark.gpu

SCHEDULES = [(16, 12, 0, 0), (8, 8, 1, 2), (4096, 1, 2, 1), (3, 5, 1, 1), (64, 4, 3, 4)]


def run_with(sc, prec, sched, nsteps=31):
    d = domain_for(sc, precision=prec)
    d.set_schedule(*sched)
    got = d.schedule()
    assert (got["tile_x"], got["tile_y"], got["wg_pred"], got["wg_corr"]) == sched
    assert got["grid_pred"] % 8 == 0 and got["grid_corr"] % 8 == 0
    L = bgf.Loop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    d.init_dt(L)
    d.run(L, nsteps)
    return d.get_state(), L


@pytest.mark.parametrize("prec", ["fp64", "fp32_fast", "fp32_comp"])
def test_domain_results_do_not_depend_on_schedule(prec):
    sc = scenarios.rain_terrain(n=160, dx=5.0, amp=12.0, rain=60.0, manning=0.03, seed=5,
                                precision="fp64" if prec == "fp64" else "fp32")
    T = np.float64 if prec == "fp64" else np.float32
    m = interior(sc)
    sc["h"] = np.where(m, np.maximum(T(9.0) - sc["zb"], 0), 0).astype(T)
    ref, Lr = run_with(sc, prec, SCHEDULES[0])
    for s in SCHEDULES[1:]:
        print("case schedule domain %s %s" % (prec, s), flush=True)
        st, L = run_with(sc, prec, s)
        assert (L.totaltime, L.dt, L.dtmax) == (Lr.totaltime, Lr.dt, Lr.dtmax)
        for k in ("h", "u", "v", "zs"):
            assert np.array_equal(st[k][m], ref[k][m]), (s, k)


def test_schedule_change_between_runs_keeps_the_trajectory():
    prec = "fp32_fast"
    sc = members(prec, sizes=(96,))[0]
    d = domain_for(sc, precision=prec)
    L = bgf.Loop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    d.init_dt(L)
    for s in SCHEDULES:
        d.set_schedule(*s)
        d.run(L, 7)
    ref, Lr = solo(sc, prec, 1e9, 7 * len(SCHEDULES))
    assert (L.totaltime, L.dt) == (Lr.totaltime, Lr.dt)
    mk = interior(sc)
    st, rs = d.get_state(), ref.get_state()
    for k in ("h", "u", "v"):
        assert np.array_equal(st[k][mk], rs[k][mk]), k


def test_batch_results_do_not_depend_on_schedule():
    prec = "fp32_comp"
    scs = members(prec)
    print("case schedule batch members=%d" % len(scs), flush=True)
    common = dict(scs[0]["param"], precision=prec)
    B = bgf.Batch(common, max_blocks=sum(sc["param"]["nblkmem"] for sc in scs) + 40, max_members=8)
    ids = add_all(B, scs, [1e9] * len(scs))
    nsteps = 0
    for s in SCHEDULES[::-1]:
        B.set_schedule(*s)
        assert tuple(B.schedule().values()) == s
        B.run(6)
        nsteps += 6
    for m, sc in zip(ids, scs):
        d, L = solo(sc, prec, 1e9, nsteps)
        assert_member_equals(B, m, sc, d, L)


def test_autotune_domain_and_cache(tmp_path):
    sc = scenarios.rain_terrain(n=128, dx=5.0, amp=12.0, rain=60.0, manning=0.03, seed=3, precision="fp32")
    p = dict(sc["param"], precision="fp32_fast")
    cache = str(tmp_path / "sched.json")
    sched, trials = autotune.autotune_domain(p, sc["blocks"], sc["zb"], sc["h"], sc["u"], sc["v"], sc["cf"],
                                             steps=4, repeats=1, cache=cache, tiles=((16, 12), (8, 8)), max_wg=2)
    assert len(trials) == 2 + 2 + 2
    assert all(t["s_per_step"] > 0 for t in trials)
    best = min(trials, key=lambda t: t["s_per_step"])
    assert best["s_per_step"] <= trials[0]["s_per_step"]
    again, t2 = autotune.autotune_domain(p, sc["blocks"], sc["zb"], sc["h"], sc["u"], sc["v"], sc["cf"], cache=cache)
    assert again == sched and t2 == []
    d = domain_for(sc, precision="fp32_fast")
    d.set_schedule(**sched)
    assert {k: d.schedule()[k] for k in sched} == sched
