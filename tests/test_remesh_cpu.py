"""In-run adaptation on the CPU: the adaptation pass with coarsening (the reference's sanity rules
and block numbering, src/Adaptation.cu:157-1876) against the oracle's restatement, the properties
the rules fix, and the oracle's state remap (refine's bilinear interpolation, coarsen's means, the
bed from a raster) pinned by exact cases."""
import numpy as np
import pytest

import bg_flood_amd as bgf
from bg_flood_amd import scenarios
from oracle import pyoracle as po

KEYS = ("level", "xo", "yo", "LeftBot", "LeftTop", "RightBot", "RightTop", "TopLeft", "TopRight", "BotLeft",
        "BotRight")
BW, MW, BS = 16, 18, 324


def params(nbx, nby, dx, nblk, nblkmem, maxlevel, prec, **kw):
    p = bgf.default_param(dx=dx, delta=dx, nblk=nblk, nblkmem=nblkmem, minlevel=0, maxlevel=maxlevel,
                          precision=prec, **kw)
    op = po.make_param(dx=dx, delta=dx, nblk=nblk, nblkmem=nblkmem, minlevel=0, maxlevel=maxlevel,
                       nx=nbx * BW, ny=nby * BW, doubleprecision=1 if prec == "fp64" else 0,
                       **{k: v for k, v in kw.items() if k in ("eps", "wetdryfix")})
    return p, op


def both_passes(blocks, nblk, nblkmem, ref, co, nbx, nby, dx, maxlevel, prec):
    p, op = params(nbx, nby, dx, nblk, nblkmem, maxlevel, prec)
    lib = bgf.adapt_mesh(p, blocks, ref, co, nbx, nby)
    orc = po.adapt_mesh(op, blocks, ref, co, 4 * nblkmem + 16)
    return lib, orc


def assert_same_mesh(lib, orc):
    (lb, ln, lm, lr), (ob, on, om, orr) = lib, orc
    assert (ln, lm) == (on, om)
    assert np.array_equal(lb["active"][:ln], ob["active"][:on])
    live = np.asarray(lb["active"][:ln])
    for k in KEYS:
        assert np.array_equal(lb[k][live], ob[k][live]), k
    assert np.array_equal(lr[live], orr[live])
    assert np.all(lr[:, 0][np.setdiff1d(np.arange(lm), live)] == -1)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_adaptation_with_coarsening_matches_the_oracle(prec, seed):
    rng = np.random.default_rng(seed)
    nbx, nby, dx, maxlevel = 5, 4, 7.3, 3
    blocks, nblk, nblkmem = scenarios.adapt_mesh(nbx, nby, dx, 2, lambda x, y: np.hypot(x - 200, y - 150) < 90,
                                                 precision=prec)
    for it in range(6):
        act = np.asarray(blocks["active"][:nblk])
        ref = np.zeros(nblkmem, np.uint8)
        ref[act] = rng.random(nblk) < 0.12
        co = np.zeros(nblkmem, np.uint8)
        co[act] = (1 - ref[act]) * (rng.random(nblk) < 0.9)
        lib, orc = both_passes(blocks, nblk, nblkmem, ref, co, nbx, nby, dx, maxlevel, prec)
        assert_same_mesh(lib, orc)
        blocks, nblk, nblkmem, remap = lib
        kinds = remap[np.asarray(blocks["active"][:nblk]), 0]
        print("pass", it, nblk, np.bincount(kinds, minlength=3))


def test_refine_then_coarsen_restores_the_mesh():
    nbx, nby, dx = 4, 3, 10.0
    blocks = bgf.mesh_uniform(nbx, nby, dx, 0, True)
    nblk, nblkmem = 12, scenarios.nblkmem_for(12)
    ref = np.zeros(nblkmem, np.uint8)
    ref[[1, 6, 7]] = 1
    p, _ = params(nbx, nby, dx, nblk, nblkmem, 2, "fp64")
    b1, n1, m1, r1 = bgf.adapt_mesh(p, blocks, ref, None, nbx, nby)
    assert n1 == 12 + 9 and np.all(r1[[1, 6, 7], 0] == 1)
    # coarsen everything: exactly the three refined quads merge back
    p, _ = params(nbx, nby, dx, n1, m1, 2, "fp64")
    co = np.zeros(m1, np.uint8)
    co[b1["active"][:n1]] = 1
    b2, n2, m2, r2 = bgf.adapt_mesh(p, b1, np.zeros(m1, np.uint8), co, nbx, nby)
    assert n2 == 12 and np.array_equal(b2["active"][:n2], blocks["active"][:12])
    for k in KEYS:
        assert np.array_equal(np.asarray(b2[k])[:12], np.asarray(blocks[k])[:12]), k
    assert sorted(np.flatnonzero(r2[:, 0] == 2)) == [1, 6, 7]
    assert np.array_equal(r2[1, 1:], [1, b1["RightBot"][1], b1["TopLeft"][1], b1["TopLeft"][b1["RightBot"][1]]])


def test_coarsening_rules():
    nbx, nby, dx = 4, 4, 10.0
    blocks = bgf.mesh_uniform(nbx, nby, dx, 0, True)
    nblk, nblkmem = 16, scenarios.nblkmem_for(16)
    ref = np.zeros(nblkmem, np.uint8)
    ref[[0, 5, 10]] = 1
    p, _ = params(nbx, nby, dx, nblk, nblkmem, 2, "fp64")
    b1, n1, m1, _ = bgf.adapt_mesh(p, blocks, ref, None, nbx, nby)
    act = np.asarray(b1["active"][:n1])
    p, _ = params(nbx, nby, dx, n1, m1, 2, "fp64")
    zero = np.zeros(m1, np.uint8)
    # without coarsen flags nothing coarsens (the reference cancels coarsening)
    _, n, _, r = bgf.adapt_mesh(p, b1, zero, None, nbx, nby)
    assert n == n1 and not np.any(r[act, 0] == 2)
    # one sibling of quad 0 does not ask: quad 0 stays; quads 5 and 10 merge
    co = zero.copy()
    co[act] = 1
    co[b1["TopLeft"][0]] = 0
    _, n, _, r = bgf.adapt_mesh(p, b1, zero, co, nbx, nby)
    assert sorted(np.flatnonzero(r[:, 0] == 2)) == [5, 10] and n == n1 - 6
    # a refining neighbour keeps the quads beside it (checkneighbourrefine, all eight neighbours)
    ref2 = zero.copy()
    ref2[b1["RightBot"][b1["RightBot"][5]]] = 1  # level-0 block right of quad 5
    co = zero.copy()
    co[act] = 1 - ref2[act]
    _, n, _, r = bgf.adapt_mesh(p, b1, ref2, co, nbx, nby)
    assert 5 not in np.flatnonzero(r[:, 0] == 2)
    # level-0 blocks never coarsen (minlevel)
    assert not np.any(r[act[b1["level"][act] == 0], 0] == 2)


def oracle_for(blocks, nblk, nblkmem, nbx, nby, dx, prec, maxlevel=2, eps=1e-4):
    _, op = params(nbx, nby, dx, nblk, nblkmem, maxlevel, prec, eps=eps)
    o = po.Oracle(op, blocks)
    return o, op


def state_arrays(blocks, nblkmem, dx, T, fh, fu, fv, fzb, cf=0.01):
    X, Y = scenarios.cell_coords_levels(blocks, nblkmem, 0.0, 0.0, dx, T)
    zb = fzb(X, Y).astype(T)
    h = fh(X, Y).astype(T)
    u = fu(X, Y).astype(T)
    v = fv(X, Y).astype(T)
    c = np.full(X.shape, cf, T)
    return [a.reshape(-1) for a in (zb, h, u, v, c)]


def new_param(op, nblk, nblkmem):
    q = po.make_param(**{f: getattr(op, f) for f, _ in po.OrcParam._fields_})
    q.nblk, q.nblkmem = nblk, nblkmem
    return q


def remesh_case(prec, fh, fu, fv, fzb, ref_ids=(5, 6), coarsen=True, bed=None, eps=1e-4):
    """A 2-level mesh; refine blocks ref_ids and coarsen every other quad; the oracle state before and
    after, with the cell centres of the new mesh."""
    T = np.float64 if prec == "fp64" else np.float32
    nbx, nby, dx = 4, 4, 8.0
    blocks = bgf.mesh_uniform(nbx, nby, dx, 0, prec == "fp64")
    nblk, nblkmem = 16, scenarios.nblkmem_for(16)
    ref = np.zeros(nblkmem, np.uint8)
    ref[[0, 10, 15]] = 1
    p, _ = params(nbx, nby, dx, nblk, nblkmem, 2, prec)
    b1, n1, m1, _ = bgf.adapt_mesh(p, blocks, ref, None, nbx, nby)
    o, op = oracle_for(b1, n1, m1, nbx, nby, dx, prec, eps=eps)
    o.set_state(*state_arrays(b1, m1, dx, T, fh, fu, fv, fzb))
    act = np.asarray(b1["active"][:n1])
    r = np.zeros(m1, np.uint8)
    r[list(ref_ids)] = 1
    co = None
    if coarsen:
        co = np.zeros(m1, np.uint8)
        co[act] = 1 - r[act]
    b2, n2, m2, remap = bgf.adapt_mesh(bgf.default_param(dx=dx, delta=dx, nblk=n1, nblkmem=m1, minlevel=0,
                                                          maxlevel=2, precision=prec), b1, r, co, nbx, nby)
    o.remesh(new_param(op, n2, m2), b2, remap, bed=bed)
    X, Y = scenarios.cell_coords_levels(b2, m2, 0.0, 0.0, dx, np.float64)
    return o, b2, n2, m2, remap, X.reshape(-1), Y.reshape(-1)


def interior_of(blocks, nblk, nblkmem, kinds=None, remap=None):
    m = bgf.interior_mask(nblkmem, blocks["active"][:nblk]).reshape(nblkmem, MW, MW)
    if kinds is not None:
        sel = np.isin(remap[:, 0], kinds)
        m = m & sel[:, None, None]
    return m.reshape(-1)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_constant_state_stays_constant(prec):
    o, b2, n2, m2, remap, X, Y = remesh_case(prec, lambda x, y: 1.5 + 0 * x, lambda x, y: 0.25 + 0 * x,
                                             lambda x, y: -0.5 + 0 * x, lambda x, y: -4.0 + 0 * x)
    m = interior_of(b2, n2, m2)
    assert np.any(remap[b2["active"][:n2], 0] == 1) and np.any(remap[b2["active"][:n2], 0] == 2)
    assert np.all(o.get("h")[m] == 1.5) and np.all(o.get("u")[m] == 0.25) and np.all(o.get("v")[m] == -0.5)
    assert np.all(o.get("zb")[m] == -4.0) and np.all(o.get("zs")[m] == -2.5)


def test_linear_fields_are_reproduced():
    """Bilinear refinement reproduces a linear field at every child cell whose stencil (halo included)
    carries the field, and the 2x2 mean gives the coarse cell-centre value."""
    a, bx, by = 3.0, 0.0625, -0.03125
    f = lambda x, y: a + bx * x + by * y
    o, b2, n2, m2, remap, X, Y = remesh_case("fp64", f, lambda x, y: 0.5 * f(x, y) - 1, lambda x, y: 0 * x,
                                             lambda x, y: -10.0 + 0 * x, ref_ids=(5, 6, 9))
    want = f(X, Y)
    h = o.get("h")
    for kind in (1, 2):
        m = interior_of(b2, n2, m2, [kind], remap)
        assert m.any()
        assert np.max(np.abs(h[m] - want[m])) < 1e-12, kind
        assert np.max(np.abs(o.get("u")[m] - (0.5 * want[m] - 1))) < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lake_at_rest_survives_with_a_bed_raster(prec):
    """Still water over a rough bed sampled from a raster: after refining and coarsening, every wet
    new cell keeps zs = still level (h = zs - zb) and zero velocity; the bed of new cells is the
    raster's bilinear interpolation (children) or block mean (coarsened cells)."""
    rng = np.random.default_rng(3)
    nx, ny, rdx = 129, 129, 4.0
    raster = (-3.0 + rng.random((ny, nx))).astype(np.float32)
    gx = np.arange(nx) * rdx
    gy = np.arange(ny) * rdx

    def zb_fn(x, y):
        i = np.clip((x / rdx).astype(int), 0, nx - 1)
        j = np.clip((y / rdx).astype(int), 0, ny - 1)
        return raster[j, i].astype(np.float64)

    T = np.float64 if prec == "fp64" else np.float32
    o, b2, n2, m2, remap, X, Y = remesh_case(prec, lambda x, y: 1.0 - zb_fn(x, y), lambda x, y: 0 * x,
                                             lambda x, y: 0 * x, zb_fn, bed=(raster, 0.0, 0.0, rdx))
    m = interior_of(b2, n2, m2, [1, 2], remap)
    zs, h, zb = o.get("zs"), o.get("h"), o.get("zb")
    tol = 1e-12 if prec == "fp64" else 4e-7
    assert np.max(np.abs(zs[m] - 1.0)) <= tol
    assert np.all(o.get("u")[m] == 0) and np.all(o.get("v")[m] == 0)
    # children (level 1, dx 4 = raster spacing): bilinear of the raster at the cell centre
    kids = interior_of(b2, n2, m2, [1], remap)
    from scipy.interpolate import RegularGridInterpolator
    ip = RegularGridInterpolator((gy, gx), raster.astype(np.float64))
    want = ip(np.stack([Y[kids], X[kids]], axis=1))
    assert np.max(np.abs(zb[kids] - want)) <= (1e-12 if prec == "fp64" else 2e-6)
    # coarsened cells (level 0, dx 8): the mean of the raster cells the cell covers
    co = interior_of(b2, n2, m2, [2], remap)
    xs, ys = X[co], Y[co]
    for k in range(0, xs.size, 97):
        i0 = int(np.floor((xs[k] - 4.0) / rdx)); i1 = int(np.floor((xs[k] + 4.0) / rdx))
        j0 = int(np.floor((ys[k] - 4.0) / rdx)); j1 = int(np.floor((ys[k] + 4.0) / rdx))
        want = raster[max(j0, 0):min(j1, ny - 1) + 1, max(i0, 0):min(i1, nx - 1) + 1].astype(np.float64).mean()
        assert abs(zb[co][k] - want) <= (1e-12 if prec == "fp64" else 2e-6)


def test_coarsening_conserves_volume_without_a_raster():
    field = lambda x, y: 2.0 + np.sin(0.05 * x) * np.cos(0.07 * y)
    o, b2, n2, m2, remap, X, Y = remesh_case("fp64", field, lambda x, y: 0 * x, lambda x, y: 0 * x,
                                             lambda x, y: -5.0 + 0 * x, ref_ids=())
    assert np.any(remap[:, 0] == 2) and not np.any(remap[:, 0] == 1)
    # every cell deep (h > eps): the fix-up keeps h = zs - zb = mean(h) with a flat bed
    T = np.float64
    nbx, nby, dx = 4, 4, 8.0
    blocks = bgf.mesh_uniform(nbx, nby, dx, 0, True)
    ref = np.zeros(17, np.uint8)
    ref[[0, 10, 15]] = 1
    b1, n1, m1, _ = bgf.adapt_mesh(bgf.default_param(dx=dx, delta=dx, nblk=16, nblkmem=17, minlevel=0,
                                                     maxlevel=2), blocks, ref, None, nbx, nby)
    o1, _ = oracle_for(b1, n1, m1, nbx, nby, dx, "fp64")
    o1.set_state(*state_arrays(b1, m1, dx, T, field, lambda x, y: 0 * x, lambda x, y: 0 * x,
                               lambda x, y: -5.0 + 0 * x))
    assert abs(o.volume() - o1.volume()) <= 1e-12 * o1.volume()


def test_lake_at_rest_away_from_the_shore_on_an_island_mesh():
    """Still water around an island on a 3-level mesh, refined where the bed crosses the still level
    and coarsened elsewhere, bed re-sampled from a raster: every cell deeper than 1 m keeps zs = 0
    exactly, including child cells at corners beside coarser diagonal blocks (the corner is
    extrapolated from the block's own tile there)."""
    from tests.test_gpu_remesh import scenario, raster, NBX, NBY
    from tests._util import oracle_for, orc_param_with
    sc = scenario("fp64")
    sc["h"] = np.maximum(0.0 - sc["zb"], 0.0)
    o = oracle_for(sc, precision="fp64")
    p, blocks = sc["param"], sc["blocks"]
    zb = sc["zb"].reshape(-1, MW, MW)[:, 1:MW - 1, 1:MW - 1]
    act = np.asarray(blocks["active"][:p["nblk"]])
    ref = np.zeros(p["nblkmem"], np.uint8)
    ref[act] = ((zb[act] >= -0.5) & (zb[act] <= 0.5)).any(axis=(1, 2))
    co = np.zeros(p["nblkmem"], np.uint8)
    co[act] = 1 - ref[act]
    b2, nb, nm, remap = bgf.adapt_mesh(p, blocks, ref, co, NBX, NBY)
    assert np.any(remap[b2["active"][:nb], 0] == 1)
    o.remesh(orc_param_with(o.p, nblk=nb, nblkmem=nm), b2, remap, bed=raster())
    m = interior_of(b2, nb, nm)
    zs, z = o.get("zs")[m], o.get("zb")[m]
    deep = z < -1.0
    assert deep.sum() > 1000 and np.max(np.abs(zs[deep])) == 0.0
