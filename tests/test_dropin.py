"""The C++ drop-in (include/bgflood_amd.hpp) compiles against BG_Flood-shaped types with the host
compiler alone, links the library, and (on a GPU) steps exactly like the Python binding."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "bg_flood_amd")


def build(tmp_path):
    exe = str(tmp_path / "dropin")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp"), "-L", LIBDIR, "-lbgflood_amd",
                    "-Wl,-rpath," + LIBDIR, "-o", exe], check=True)
    return exe


def test_dropin_compiles_and_reports_missing_device(tmp_path):
    exe = build(tmp_path)
    import bg_flood_amd as bgf
    if bgf.lib().bgf_device_count() > 0:
        pytest.skip("a GPU is visible: covered by the gpu test")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 3 and "no device" in r.stdout


@pytest.mark.gpu
def test_dropin_steps_like_the_python_binding(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = [float(x) for x in r.stdout.split()]
    import bg_flood_amd as bgf
    from bg_flood_amd import scenarios
    p = bgf.default_param(dx=10.0, delta=10.0, nblk=4, nblkmem=5, precision="fp64", aoibnd=0, frictionmodel=0)
    blocks = bgf.mesh_uniform(2, 2, 10.0, 0, True)
    S = 5 * 324
    h = np.zeros(S)
    for ib in range(4):
        for j in range(16):
            for i in range(16):
                x = blocks["xo"][ib] + i * 10.0 - 160.0
                y = blocks["yo"][ib] + j * 10.0 - 160.0
                h[(i + 1) + (j + 1) * 18 + ib * 324] = 10.0 + np.exp(-(x * x + y * y) / (2.0 * 60.0 * 60.0))
    d = bgf.Domain(p, blocks)
    d.set_state(np.full(S, -10.0), h, np.zeros(S), np.zeros(S), np.zeros(S))
    L = bgf.Loop(nextoutputtime=1e9, dt=0, dtmax=0, totaltime=0)
    d.init_dt(L)
    dts = []
    for _ in range(5):
        d.step(L)
        L.totaltime += L.dt
        dts.append(L.dt)
    assert vals[:5] == dts
    assert vals[5] == d.get_state()["h"][(5 + 1) + (7 + 1) * 18]
