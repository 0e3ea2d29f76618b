"""Multi-level meshes on the CPU: the quadtree mesh builder against a brute-force geometric
neighbour search, and pins of the oracle's coarse/fine branches (halo fill, linear prolongation,
flux restriction, wet/dry exchanges) by invariants the mathematics fixes."""
import numpy as np
import pytest

import bg_flood_amd as bgf
from bg_flood_amd import scenarios
from oracle import pyoracle as po
from tests._util import oracle_for, interior

BW = 16


def brute_force_neighbours(leaves, nbx, nby):
    """Neighbours by geometry: the leaf containing a point just across each side, near each end
    of the side (same convention as the reference's BlockP: first half = bottom/left)."""
    L = max(l for l, _, _ in leaves)
    box = []
    for l, x, y in leaves:
        s = 1 << (L - l)  # size in finest-level block units
        box.append((x * s, y * s, s))

    def at(px, py):  # px, py in half finest-block units (odd values avoid edges)
        for k, (x0, y0, s) in enumerate(box):
            if 2 * x0 < px < 2 * (x0 + s) and 2 * y0 < py < 2 * (y0 + s):
                return k
        return None

    W, H = nbx << L, nby << L
    out = {k: [] for k in ("LeftBot", "LeftTop", "RightBot", "RightTop", "BotLeft", "BotRight", "TopLeft", "TopRight")}
    for k, (x0, y0, s) in enumerate(box):
        def nb(px, py):
            if px < 0 or py < 0 or px > 2 * W or py > 2 * H:
                return k
            return at(px, py)
        lo, hi_ = 2 * x0 + 1, 2 * (x0 + s) - 1
        blo, bhi = 2 * y0 + 1, 2 * (y0 + s) - 1
        out["LeftBot"].append(nb(2 * x0 - 1, blo)); out["LeftTop"].append(nb(2 * x0 - 1, bhi))
        out["RightBot"].append(nb(2 * (x0 + s) + 1, blo)); out["RightTop"].append(nb(2 * (x0 + s) + 1, bhi))
        out["BotLeft"].append(nb(lo, 2 * y0 - 1)); out["BotRight"].append(nb(hi_, 2 * y0 - 1))
        out["TopLeft"].append(nb(lo, 2 * (y0 + s) + 1)); out["TopRight"].append(nb(hi_, 2 * (y0 + s) + 1))
    return out


@pytest.mark.parametrize("maxlevel,nb", [(1, 3), (2, 4), (3, 5)])
def test_mesh_from_leaves_matches_geometry(maxlevel, nb):
    rng = np.random.default_rng(maxlevel)
    flags = {}
    leaves = scenarios.quadtree_leaves(nb, nb + 1, maxlevel,
                                       lambda l, x, y: flags.setdefault((l, x, y), rng.random() < 0.45))
    b = bgf.mesh_from_leaves(leaves, nb, nb + 1, 10.0)
    ref = brute_force_neighbours(leaves, nb, nb + 1)
    for k, v in ref.items():
        assert list(b[k]) == v, k
    assert list(b["active"]) == list(range(len(leaves)))
    assert list(b["level"]) == [l for l, _, _ in leaves]
    # 2:1 balance
    for k in ref:
        for i, n in enumerate(b[k]):
            assert abs(b["level"][n] - b["level"][i]) <= 1


def test_mesh_from_leaves_rejects_bad_trees():
    with pytest.raises(bgf.BgfError):  # a hole
        bgf.mesh_from_leaves([(0, 0, 0)], 2, 1, 1.0)
    with pytest.raises(bgf.BgfError):  # overlap
        bgf.mesh_from_leaves([(0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 0, 1), (1, 1, 1)], 1, 1, 1.0)
    with pytest.raises(bgf.BgfError):  # level jump of two
        lv = [(0, 1, 0)] + [(1, 0, 0), (1, 0, 1)] + [(2, 2, y) for y in range(4)] + [(2, 3, y) for y in range(4)]
        bgf.mesh_from_leaves(lv, 2, 1, 1.0)


def _fine_halo_positions(sc):
    """(cell index, x, y) of the halo cells of blocks beside a coarser block."""
    b, p = sc["blocks"], sc["param"]
    out = []
    for ib in range(p["nblk"]):
        lev = b["level"][ib]
        d = p["dx"] / (1 << lev)
        x0, y0 = b["xo"][ib], b["yo"][ib]
        for side, nbk in (("L", "LeftBot"), ("R", "RightBot"), ("B", "BotLeft"), ("T", "TopLeft")):
            if b["level"][b[nbk][ib]] >= lev:
                continue
            for j in range(BW):
                i, jj = {"L": (-1, j), "R": (BW, j), "B": (j, -1), "T": (j, BW)}[side]
                out.append(((i + 1) + (jj + 1) * 18 + ib * 324, x0 + i * d, y0 + jj * d))
    return out


def test_linear_fields_are_reproduced_at_coarse_fine_halos():
    """fillHalo's 1/3-1/6-1/2 weights (on zb) and refine_linear (on u, v) reproduce linear fields
    at fine halo cells beside coarser blocks; the fine-to-coarse 2x2 mean does at coarse halos."""
    sc = scenarios.amr_basin(nbx=4, nby=4, maxlevel=2, amp=0.0, precision="fp64")
    X, Y = scenarios.cell_coords_levels(sc["blocks"], sc["param"]["nblkmem"], 0.0, 0.0, sc["param"]["dx"],
                                        np.float64)
    m = interior(sc)
    a, bx, by = 0.3, 2e-4, -3e-4
    sc["u"] = np.where(m, a + bx * X + by * Y, 0.0)
    sc["v"] = np.where(m, -a + by * X - bx * Y, 0.0)
    sc["zb"] = np.where(m, -10.0 + 1e-3 * X + 2e-3 * Y, 0.0)
    sc["h"] = np.where(m, 10.0 - (1e-3 * X + 2e-3 * Y) + 0.5, 0.0)  # flat free surface
    o = oracle_for(sc)
    zb = o.get("zb")
    pos = _fine_halo_positions(sc)
    assert len(pos) > 50
    for c, x, y in pos:
        assert zb[c] == pytest.approx(-10.0 + 1e-3 * x + 2e-3 * y, abs=1e-12)
    loop = po.OrcLoop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    o.init_loop(loop)
    o.step(loop)  # the evolv halos now hold the stage-1 exchange of the initial state
    u, v = o.get("u"), o.get("v")
    for c, x, y in pos:
        assert u[c] == pytest.approx(a + bx * x + by * y, abs=1e-12)
        assert v[c] == pytest.approx(-a + by * x - bx * y, abs=1e-12)


@pytest.mark.parametrize("island", [False, True])
def test_amr_mass_conservation_with_walls(island):
    """Walls and flux restriction: volume changes only by round-off (a coarse face flux that is
    not the mean of its fine faces would leak mass at every level change)."""
    sc = scenarios.amr_basin(island=island, precision="fp64")
    o = oracle_for(sc)
    loop = po.OrcLoop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    o.init_loop(loop)
    v0 = o.volume()
    for _ in range(60):
        o.step(loop)
        loop.totaltime += loop.dt
    assert abs(o.volume() - v0) / v0 < 1e-12


@pytest.mark.parametrize("level", [0.5, -7.0, -8.0])
def test_amr_lake_at_rest_on_a_plane(level):
    """Still water over a planar bed stays at rest across coarse/fine interfaces, with the
    shoreline crossing them for the lower levels: every exchange (2x2 mean, 1/3-1/6-1/2 weights,
    refine_linear, wet/dry restriction and prolongation) is exact for linear fields.  (Over a
    non-planar bed the interpolated bed halo and the limited linear reconstruction of h disagree
    beside coarser blocks, and the reference's engine 2 does not keep the lake at rest there.)"""
    sc = scenarios.amr_basin(amp=0.0, precision="fp64")
    p = sc["param"]
    X, Y = scenarios.cell_coords_levels(sc["blocks"], p["nblkmem"], 0.0, 0.0, p["dx"], np.float64)
    m = interior(sc)
    sc["zb"] = np.where(m, -10.0 + 0.004 * X + 0.002 * Y, 0.0)
    sc["h"] = np.where(m, np.maximum(level - sc["zb"], 0.0), 0.0)
    o = oracle_for(sc)
    loop = po.OrcLoop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    o.init_loop(loop)
    for _ in range(40):
        o.step(loop)
        loop.totaltime += loop.dt
    h, u, v, zb = o.get("h")[m], o.get("u")[m], o.get("v")[m], o.get("zb")[m]
    wet = h > 1e-4
    assert np.abs(u).max() < 1e-12 and np.abs(v).max() < 1e-12
    assert np.abs((h + zb)[wet] - level).max() < 1e-12


def test_amr_hump_is_symmetric():
    """The centred hump on a mesh refined symmetrically about the centre keeps the mirror
    symmetries x -> L-x and y -> L-y (a wrong sign or index on one side breaks them)."""
    sc = scenarios.amr_basin(nbx=4, nby=4, maxlevel=2, precision="fp64")
    o = oracle_for(sc)
    loop = po.OrcLoop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    o.init_loop(loop)
    for _ in range(40):
        o.step(loop)
        loop.totaltime += loop.dt
    p = sc["param"]
    X, Y = scenarios.cell_coords_levels(sc["blocks"], p["nblkmem"], 0.0, 0.0, p["dx"], np.float64)
    m = interior(sc)
    h = o.get("h")
    L = 4 * BW * p["dx"]
    key = {(round(x, 6), round(y, 6)): i for i, x, y in zip(np.nonzero(m)[0], X[m], Y[m])}
    n = 0
    for (x, y), i in key.items():
        for xm, ym in ((L - x, y), (x, L - y), (y, x)):
            j = key.get((round(xm, 6), round(ym, 6)))
            if j is not None:
                assert h[i] == pytest.approx(h[j], abs=1e-11)
                n += 1
    assert n > 1000
