"""Refinement passes (A19) on the CPU: the library's pass with the reference's block numbering
against the oracle's restatement, bit for bit, and the properties the reference's algorithm fixes."""
import numpy as np
import pytest

import bg_flood_amd as bgf
from bg_flood_amd import scenarios
from oracle import pyoracle as po

KEYS = ("active", "level", "xo", "yo", "LeftBot", "LeftTop", "RightBot", "RightTop", "TopLeft", "TopRight",
        "BotLeft", "BotRight")


def one_pass(blocks, nblk, nblkmem, flags, nbx, nby, dx, maxlevel, prec):
    p = bgf.default_param(dx=dx, delta=dx, nblk=nblk, nblkmem=nblkmem, maxlevel=maxlevel, precision=prec)
    lib = bgf.adapt_refine(p, blocks, flags, nbx, nby)
    op = po.make_param(dx=dx, delta=dx, nblk=nblk, nblkmem=nblkmem, maxlevel=maxlevel, nx=nbx * 16, ny=nby * 16,
                       doubleprecision=1 if prec == "fp64" else 0)
    orc = po.adapt_refine(op, blocks, flags, 4 * nblkmem + 16)
    return lib, orc


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_refinement_passes_match_the_oracle(prec, seed):
    rng = np.random.default_rng(seed)
    nbx, nby, dx, maxlevel = 5, 4, 7.3, 3
    blocks = bgf.mesh_uniform(nbx, nby, dx, 0, prec == "fp64")
    nblk, nblkmem = nbx * nby, scenarios.nblkmem_for(nbx * nby)
    for it in range(3):
        flags = np.zeros(nblkmem, np.uint8)
        flags[np.asarray(blocks["active"][:nblk])] = rng.random(nblk) < 0.3
        (lb, ln, lm), (ob, on, om) = one_pass(blocks, nblk, nblkmem, flags, nbx, nby, dx, maxlevel, prec)
        assert (ln, lm) == (on, om)
        for k in KEYS:
            n = ln if k == "active" else lm
            a, b = lb[k][:n], ob[k][:n]
            if k in ("level", "xo", "yo") or k not in ("active",):
                live = np.asarray(lb["active"][:ln])
                a, b = a[live], b[live]
            assert np.array_equal(a, b), (it, k)
        blocks, nblk, nblkmem = lb, ln, lm


def test_refinement_pass_numbering_and_balance():
    """Refining blocks keep their index (bottom-left child), the other children take the lowest free
    slots in refining order and are appended to the active list; the 2:1 cascade refines coarser
    neighbours; nothing refines beyond maxlevel."""
    nbx, nby, dx = 4, 4, 10.0
    blocks = bgf.mesh_uniform(nbx, nby, dx, 0, True)
    nblk, nblkmem = 16, scenarios.nblkmem_for(16)
    flags = np.zeros(nblkmem, np.uint8)
    flags[5] = 1
    b1, n1, m1 = bgf.adapt_refine(bgf.default_param(dx=dx, nblk=nblk, nblkmem=nblkmem, maxlevel=2), blocks, flags,
                                  nbx, nby)
    assert n1 == 19 and list(b1["active"][:16]) == list(range(16))
    free = [k for k in range(m1) if k >= 16]
    assert list(b1["active"][16:19]) == free[:3]
    assert b1["level"][5] == 1 and all(b1["level"][k] == 1 for k in free[:3])
    # refine a level-1 child at the corner: its coarser level-0 neighbours must refine too
    flags = np.zeros(m1, np.uint8)
    flags[free[2]] = 1  # top-right child of block 5
    b2, n2, m2 = bgf.adapt_refine(bgf.default_param(dx=dx, nblk=n1, nblkmem=m1, maxlevel=2), b1, flags, nbx, nby)
    lev = b2["level"]
    act = b2["active"][:n2]
    for k in ("LeftBot", "RightBot", "TopLeft", "BotLeft", "LeftTop", "RightTop", "TopRight", "BotRight"):
        for ib in act:
            assert abs(lev[b2[k][ib]] - lev[ib]) <= 1
    assert max(lev[act]) == 2
    flags = np.ones(m2, np.uint8)
    b3, n3, m3 = bgf.adapt_refine(bgf.default_param(dx=dx, nblk=n2, nblkmem=m2, maxlevel=2), b2, flags, nbx, nby)
    assert max(b3["level"][b3["active"][:n3]]) == 2


def test_setup_adaptation_loop_reaches_a_fixed_point():
    """The set-up loop refines around a circle to maxlevel and stops when no block is added; the
    mesh runs on the oracle."""
    nbx, nby, dx = 6, 6, 5.0
    blocks, nblk, nblkmem = scenarios.adapt_mesh(nbx, nby, dx, 2, lambda x, y: np.hypot(x - 240, y - 240) < 60)
    lev = blocks["level"][blocks["active"][:nblk]]
    assert lev.max() == 2 and lev.min() == 0
    # every cell within the circle sits in a level-2 block
    hit = scenarios.block_any(blocks, nblk, nblkmem, dx, lambda x, y: np.hypot(x - 240, y - 240) < 60)
    assert all(blocks["level"][ib] == 2 for ib in blocks["active"][:nblk] if hit[ib])
