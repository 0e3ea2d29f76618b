"""Pins of the oracle's primitives against closed forms and textbook definitions (CPU)."""
import ctypes
import itertools
import math

import numpy as np
import pytest

from oracle import pyoracle as po

L = po.lib()


def textbook_minmod(theta, s0, s1, s2):
    """Generalised minmod limiter: minmod(theta*(s1-s0), (s2-s0)/2, theta*(s2-s1))."""
    a, b, c = theta * (s1 - s0), (s2 - s0) / 2.0, theta * (s2 - s1)
    if a > 0 and b > 0 and c > 0:
        return min(a, b, c)
    if a < 0 and b < 0 and c < 0:
        return max(a, b, c)
    return 0.0


@pytest.mark.parametrize("theta", [1.0, 1.3, 2.0])
def test_minmod2_exhaustive_small_integers(theta):
    for s0, s1, s2 in itertools.product(range(-3, 4), repeat=3):
        got = L.orc_minmod2_d(theta, s0, s1, s2)
        assert got == textbook_minmod(theta, float(s0), float(s1), float(s2)), (theta, s0, s1, s2)


def test_minmod2_random_and_special_cases():
    rng = np.random.default_rng(0)
    for _ in range(20000):
        th = 1.0 + rng.random()
        s = rng.normal(size=3) * 10.0 ** rng.integers(-6, 4)
        assert L.orc_minmod2_d(th, *s) == pytest.approx(textbook_minmod(th, *s), rel=1e-15, abs=0)
    # theta = 1 is plain minmod of the one-sided slopes (central slope never smaller)
    assert L.orc_minmod2_d(1.0, 0.0, 1.0, 3.0) == 1.0
    # theta = 2 (superbee-like) limits by the central difference here
    assert L.orc_minmod2_d(2.0, 0.0, 1.0, 3.0) == 1.5
    # an extremum gives zero slope
    assert L.orc_minmod2_d(1.3, 0.0, 2.0, 1.0) == 0.0
    # float instantiation agrees on exactly representable cases
    assert L.orc_minmod2_f(1.5, 0.0, 1.0, 4.0) == 1.5


def kurg(hp, hm, up, um, g=9.81, delta=1.0, CFL=0.5, epsi=np.finfo(float).eps):
    fh = ctypes.c_double()
    fu = ctypes.c_double()
    dt = L.orc_kurgsolver_d(g, delta, epsi, CFL, 1.0, 1.0, hp, hm, up, um, ctypes.byref(fh), ctypes.byref(fu))
    return fh.value, fu.value, dt


@pytest.mark.parametrize("h", [1e-3, 0.5, 10.0, 4000.0])
def test_kurgsolver_rest_state(h):
    fh, fu, dt = kurg(h, h, 0.0, 0.0)
    assert fh == 0.0
    assert fu == pytest.approx(9.81 * h * h / 2.0, rel=1e-14)
    assert dt == pytest.approx(0.5 * 1.0 / math.sqrt(9.81 * h), rel=1e-14)


@pytest.mark.parametrize("h,u", [(1.0, 5.0), (0.1, 2.0), (2.0, 30.0)])
def test_kurgsolver_supercritical_is_upwind(h, u):
    # u > sqrt(g h) on both sides: the central-upwind flux reduces to the left ("minus") state's flux
    hm, hp = h, 0.7 * h
    fh, fu, dt = kurg(hp, hm, u, u)
    assert fh == pytest.approx(hm * u, rel=1e-13)
    assert fu == pytest.approx(hm * u * u + 9.81 * hm * hm / 2.0, rel=1e-13)
    # and mirrored to the right-going/left-going case
    fh2, fu2, _ = kurg(hm, hp, -u, -u)
    assert fh2 == pytest.approx(-hm * u, rel=1e-13)
    assert fu2 == pytest.approx(hm * u * u + 9.81 * hm * hm / 2.0, rel=1e-13)


def test_kurgsolver_dry_state():
    eps = np.finfo(float).eps
    fh, fu, dt = kurg(0.0, 0.0, 0.0, 0.0)
    assert fh == 0.0 and fu == 0.0 and dt == 1.0 / eps


def test_kurgsolver_mirror_symmetry():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        hp, hm = rng.random(2) * 5
        up, um = rng.normal(size=2) * 3
        fh, fu, dt = kurg(hp, hm, up, um)
        fh2, fu2, dt2 = kurg(hm, hp, -um, -up)
        assert fh == pytest.approx(-fh2, rel=1e-12, abs=1e-12)
        assert fu == pytest.approx(fu2, rel=1e-12, abs=1e-12)
        assert dt == pytest.approx(dt2, rel=1e-15)


def test_kurgsolver_consistency_equal_states():
    # equal states give the physical flux (h u, h u^2 + g h^2/2) for any velocity
    for h, u in [(1.0, 0.3), (3.0, -1.2), (0.2, 2.5)]:
        fh, fu, _ = kurg(h, h, u, u)
        assert fh == pytest.approx(h * u, rel=1e-13)
        assert fu == pytest.approx(h * u * u + 9.81 * h * h / 2, rel=1e-13)


def test_calcres_ftoi_memloc_closed_forms():
    for lvl in range(-4, 5):
        assert L.orc_calcres_d(10.0, lvl) == 10.0 * 2.0 ** (-lvl)
    for v, e in [(0.49, 0), (0.5, 1), (1.5, 2), (-0.5, -1), (-0.49, 0), (-1.5, -2), (7.0, 7)]:
        assert L.orc_ftoi_d(v) == e
    for i, j, ib in itertools.product(range(-1, 17), range(-1, 17), range(3)):
        assert L.orc_memloc(i, j, ib) == (i + 1) + (j + 1) * 18 + ib * 324


@pytest.mark.parametrize("dbl", [True, False])
def test_timestep_reduction_semantics(dbl):
    """reductiontest (src/Testing.cu:1204-1292): planted minimum, 1.5x clamp, output landing."""
    rng = np.random.default_rng(2)
    fn = L.orc_timestep_d if dbl else L.orc_timestep_f
    eps = np.finfo(np.float64 if dbl else np.float32).eps
    for _ in range(1000):
        m = float(rng.random()) if dbl else float(np.float32(rng.random()))
        T = np.float64 if dbl else np.float32
        loop = po.OrcLoop(nextoutputtime=float(T(m) * T(2.0)), dt=0.0, dtmax=float(T(m) * T(2.01)), totaltime=0.0)
        dt = fn(m, ctypes.byref(loop))
        assert abs(dt - m) < 100 * eps * max(1.0, m)
    # growth clamp: at most 1.5x the previous step
    loop = po.OrcLoop(nextoutputtime=1e9, dt=0.0, dtmax=0.1, totaltime=0.0)
    assert fn(1.0, ctypes.byref(loop)) == pytest.approx(0.15, rel=1e-6)
    # landing: the step divides the remaining time to the next output exactly
    loop = po.OrcLoop(nextoutputtime=1.0, dt=0.0, dtmax=10.0, totaltime=0.0)
    dt = fn(0.3, ctypes.byref(loop))
    assert dt == pytest.approx(0.25, rel=1e-6)
    # no landing when the output time is now (ceil(0) == 0)
    loop = po.OrcLoop(nextoutputtime=0.0, dt=0.0, dtmax=10.0, totaltime=0.0)
    assert fn(0.3, ctypes.byref(loop)) == pytest.approx(0.3, rel=1e-6)
