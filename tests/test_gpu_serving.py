"""Continuous batching on the GPU: requests that arrive while others run are admitted as slots free
up, run overlapped, and each result (maximum depth, gauge series) is bit-identical to the same
request served alone."""
import numpy as np
import pytest

from bg_flood_amd import scenarios, serving

pytestmark = pytest.mark.gpu


def rain_request(k, n, horizon):
    sc = scenarios.rain_terrain(n=n, dx=5.0, amp=8.0 + k, rain=40.0 + 10 * k, manning=0.03, seed=100 + k,
                                precision="fp32")
    cells = [(i + 1) + (j + 1) * 18 + b * 324 for i, j, b in ((3, 4, 0), (8, 8, 1), (15, 0, 2), (7, 9, 3))]
    return serving.Request(k, sc, horizon, gauges=cells, gauge_every=3)


def amr_request(k, horizon):
    sc = scenarios.amr_basin(precision="fp32", island=True, amp=1.0 + 0.2 * k)
    cells = [(i + 1) + (j + 1) * 18 + b * 324 for i, j, b in ((5, 5, 0), (9, 2, 7))]
    return serving.Request(k, sc, horizon, gauges=cells)


def make_requests():
    return [rain_request(0, 64, 120.0), rain_request(1, 96, 60.0), rain_request(2, 48, 200.0),
            amr_request(3, 8.0), rain_request(4, 80, 90.0), rain_request(5, 64, 150.0)]


def common_param(reqs):
    return dict(reqs[0].sc["param"], precision="fp32_comp")


def test_mixed_stream_matches_requests_served_alone():
    reqs = make_requests()
    arrivals = np.array([0.0, 0.0, 0.05, 0.05, 0.1, 0.12])
    # room for about two members at a time: later arrivals wait for slots
    cap = 2 * max(r.sc["param"]["nblkmem"] for r in reqs if not r.multilevel) + 8
    srv, wall = serving.serve(reqs, arrivals, common_param(reqs), max_blocks=cap, max_members=4, chunk=16,
                              time_limit=120.0)
    print("served %d requests in %.2f s" % (len(srv.done), wall), flush=True)
    assert len(srv.done) == len(reqs)
    for r in reqs:
        assert r.latency is not None and r.latency >= 0
        assert r.result["totaltime"] == pytest.approx(r.horizon, rel=1e-6)
    for r0 in reqs:
        alone = [make_requests()[r0.rid]]
        s1, _ = serving.serve(alone, np.array([0.0]), common_param(reqs), max_blocks=cap, max_members=4, chunk=16)
        a, b = r0.result, alone[0].result
        assert np.array_equal(a["maxdepth"], b["maxdepth"]), r0.rid
        assert np.array_equal(a["gauges"], b["gauges"]), r0.rid
        assert np.array_equal(a["times"], b["times"]), r0.rid
        s1.close()
    srv.close()
