"""The drawing over device readbacks (code_mpm_auxetics_amd/host/GraphicsEngine.h),
GraphicsEngine_drawGame.cpp:4-260 of the reference: per-point draw instances
of the MAIN, J2_STRESS and J2_PLASTICSTRAIN canvases (GraphicsEngine.h:77-82).

* CPU: the colour and size rules against the reference's formulas, and the
  optional OpenGL path (MPM_WITH_GL) compiled and linked against the system
  GL library (no context is created);
* GPU: a C++ program draws a world with known stresses, plastic strains and
  flags through the C++ PhysicsEngine mirror; every instance is checked
  against the same rules evaluated here on the uploaded values.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import code_mpm_auxetics_amd._abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.dirname(A.LIB_PATH)


def _gxx():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    return gxx


def _build(tmp_path, src, name, extra=()):
    exe = tmp_path / name
    subprocess.run([_gxx(), "-std=c++17", "-Wall", "-O1", "-I", ROOT, os.path.join(ROOT, src), "-o", str(exe),
                    "-L", LIBDIR, "-lmpm_auxetics", f"-Wl,-rpath,{LIBDIR}", *extra], check=True)
    return exe


def test_drawing_rules(tmp_path):
    """MAIN colour precedence (…drawGame.cpp:73-83), the blue-to-red J2 blend
    in float (:172, :257, unclamped) and the point size 0.8 V^(1/3) (:72)."""
    exe = _build(tmp_path, "tests/native/graphics_host.cpp", "graphics_host")
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"


def test_gl_path_compiles_and_links(tmp_path):
    """The instanced OpenGL renderer builds where GL headers and libGL exist
    (compile-optional: the rest of the header needs no GL)."""
    if not os.path.exists("/usr/include/GL/gl.h") or not os.path.exists("/usr/include/GL/glext.h"):
        pytest.skip("OpenGL headers not installed")
    try:
        exe = _build(tmp_path, "tests/native/graphics_gl_link.cpp", "graphics_gl",
                     ("-DMPM_WITH_GL", "-DGL_GLEXT_PROTOTYPES", "-lGL"))
    except subprocess.CalledProcessError:
        pytest.skip("no linkable GL library")
    assert subprocess.run([str(exe)]).returncode == 0


def _j2(s):
    """calculateState_J2 (ConstitutiveRelation_calculateState_J2.cpp:4-26) of Voigt rows."""
    p = (s[:, 0] + s[:, 1] + s[:, 2]) / 3.0
    d = s[:, :3] - p[:, None]
    return 0.5 * (d ** 2).sum(axis=1) + (s[:, 3:] ** 2).sum(axis=1)


def _blend(j2, mx):
    w = np.float32(j2) / np.float32(mx)
    u = np.float32(1.0) - w
    return np.stack([w, np.zeros_like(w), u, u + w], axis=-1)


@pytest.mark.gpu
def test_draw_demo_against_the_rules(tmp_path):
    exe = _build(tmp_path, "tests/host_draw_demo.cpp", "host_draw_demo")
    print("running draw demo", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    n, n_main, n_s, n_p = map(int, lines[0].split())
    # printed with 9 significant digits: each value reads back as the same float
    rows = np.array([list(map(float, l.split())) for l in lines[1:1 + n]]).astype(np.float32)
    nodes = np.array([list(map(float, l.split())) for l in lines[1 + n:]]).astype(np.float32)
    # the demo's world
    h, nc = 0.001, 8
    ids = np.arange(32)
    i, j, k = ids // 8, (ids // 2) % 4, ids % 2
    x = np.stack([(2.25 + 0.5 * i) * h, (2.25 + 0.5 * j) * h, (2.25 + 0.5 * k) * h], axis=1)
    vol = (0.125 + 0.01 * ids) * h ** 3
    s = 1.0e5 * (ids + 1)
    st = np.stack([s, -0.5 * s, 0.25 * s, 0.1 * s, -0.2 * s, 0.05 * s * (ids % 5)], axis=1)
    q = np.arange(6)
    ep = 1.0e-5 * (ids % 7)[:, None] * (q + 1) * np.where(q % 2 == 1, -1.0, 1.0)
    mark, surf, dc = (ids % 5) == 3, (ids % 4) == 1, ids == 7
    assert n == 32 and n_s == n and n_p == n
    # positions and sizes
    assert np.array_equal(rows[:, :3], x.astype(np.float32))
    assert np.allclose(rows[:, 3], (0.8 * vol ** (1.0 / 3.0)).astype(np.float32), rtol=1e-6, atol=0)
    # MAIN: later rules win
    main = np.tile([1.0, 0.0, 0.0, 1.0], (n, 1))
    main[mark] = [0.0, 1.0, 0.0, 1.0]
    main[surf] = [0.0, 0.0, 1.0, 1.0]
    main[dc] = [1.0, 1.0, 1.0, 1.0]
    assert np.array_equal(rows[:, 4:8], main)
    # J2_STRESS: normalized by the largest J2 (at least 1e-12)
    j2s = _j2(st).astype(np.float32)
    want_s = _blend(j2s, max(np.float32(1e-12), j2s.max()))
    assert np.allclose(rows[:, 8:12], want_s, rtol=1e-6, atol=1e-7)
    assert rows[np.argmax(j2s), 8] == pytest.approx(1.0, abs=1e-7)
    # J2_PLASTICSTRAIN: normalized by yield stress / Young's modulus of point 0
    j2p = _j2(ep).astype(np.float32)
    want_p = _blend(j2p, np.float32(2.0e6 / 1.0e9))
    assert np.allclose(rows[:, 12:16], want_p, rtol=1e-6, atol=1e-7)
    assert (j2p > 0).any()
    # fixed grid nodes: the bottom layer (y fixed), gray, size 1e-5
    assert n_main == n + (nc + 1) ** 2 and len(nodes) == (nc + 1) ** 2
    assert np.allclose(nodes[:, 1], 0.0)
    assert np.array_equal(nodes[:, 3:], np.tile(np.float32([1e-5, 0.5, 0.5, 0.5, 1.0]), (len(nodes), 1)))
