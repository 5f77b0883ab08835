"""Batched independent worlds in one launch per step (GPU): each world
evolves as it does alone, per-world reactions come back separately, and the
batch matches the FP64 oracle run on the same combined scene."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mpm():
    import code_mpm_auxetics_amd as m
    m.lib()
    return m


def _worlds(mpm):
    s = mpm.scene
    return [s.small_world(material="j2", speed=2.0), s.small_world(material="plastic", speed=2.0, nx=2),
            s.small_world(material="elastic", speed=2.0, cell=0.003)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_equals_separate_runs(mpm, prec):
    p = mpm.FP64 if prec == "fp64" else mpm.FP32
    ws = _worlds(mpm)
    dt = min(w.dt_max for w in ws)
    for w in ws:
        w.dt_max = dt
    sc = mpm.scene.batch_worlds(ws)
    eng = mpm.PhysicsEngine(sc, precision=p)
    eng.run_steps(80)
    g = eng.getMaterialPoints()
    gh = eng.group_reaction_history()
    total = eng.reaction_history()
    eng.close()
    assert gh.shape == (80, 3)
    assert np.allclose(gh.sum(axis=1), total, rtol=1e-12, atol=1e-12 * np.abs(total).max())
    tol = 1e-10 if prec == "fp64" else 1e-5
    for i, w in enumerate(ws):
        e = mpm.PhysicsEngine(w, precision=p)
        e.run_steps(80)
        a = e.getMaterialPoints()
        h = e.reaction_history()
        e.close()
        sel = sc.info["world"] == i
        ox, oy = sc.info["offsets"][i]
        hw = np.array(w.length) / np.array(w.cells)
        shift = np.array([ox * hw[0], oy * hw[1], 0.0])
        ex = rel_l2(g["x"][sel] - shift - w.x, a["x"] - w.x)
        es = rel_l2(g["stress"][sel], a["stress"])
        eh = rel_l2(gh[:, i], h)
        print(f"world {i}: rel L2 disp {ex:.3e} stress {es:.3e} reaction {eh:.3e}", flush=True)
        assert ex <= tol and es <= tol * 10 and eh <= tol * 100


def test_batch_matches_oracle_fp64(mpm, oracle_mod):
    ws = _worlds(mpm)
    sc = mpm.scene.batch_worlds(ws)
    x0 = sc.x.copy()
    eng = mpm.PhysicsEngine(sc, precision=mpm.FP64)
    orc = oracle_mod.Oracle(sc)
    eng.run_steps(60)
    orc.run_steps(60)
    g, o = eng.getMaterialPoints(), orc.points()
    d = sc.dc == 0
    ex, es = rel_l2((g["x"] - x0)[d], (o["x"] - x0)[d]), rel_l2(g["stress"][d], o["stress"][d])
    print(f"rel L2 disp {ex:.3e} stress {es:.3e}", flush=True)
    assert ex <= 1e-9 and es <= 1e-9
    gh = eng.group_reaction_history()
    assert np.allclose(gh.sum(axis=1), orc.reaction, rtol=1e-8, atol=1e-8 * np.abs(orc.reaction).max())
    eng.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_persistent_launch_is_bit_identical(mpm, prec):
    """The persistent launch (a resident grid of blocks pulling bricks from
    per-XCD work queues) computes every brick exactly as one block per brick
    does: the batched config-4 cells and a rebinning free-flight block give
    bit-identical states, reactions and per-cell reactions."""
    p = mpm.FP64 if prec == "fp64" else mpm.FP32
    ws = [mpm.scene.config4_worlds(n=6)[i] for i in range(6)]
    dt = min(w.dt_max for w in ws)
    for w in ws:
        w.dt_max = dt
    for sc, steps, kw in ((mpm.scene.batch_worlds(ws), 60, {}), (mpm.scene.translation_world(), 120, {"drift_check": 4})):
        out = []
        for persistent in (False, True):
            with mpm.PhysicsEngine(sc, precision=p, persistent=persistent, **kw) as eng:
                eng.run_steps(steps)
                out.append((eng.getMaterialPoints(), eng.reaction_history(), eng.group_reaction_history(),
                            eng.checks()))
        (a, ha, ga, ca), (b, hb, gb, cb) = out
        for k in ("x", "v", "F", "stress", "eps_p", "alpha"):
            assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(ha, hb) and np.array_equal(ga, gb)
        for k in ("grid_mass", "kinetic_2", "px", "py", "pz", "rebins"):
            assert ca[k] == cb[k], k
        print(f"{prec}: {sc.n} points, {steps} steps, rebins {cb['rebins']}", flush=True)
