"""GPU readbacks beyond the point state, through the C ABI:

* per-point J2 of the stress and of the plastic strain (MPM_F_J2), the
  quantities GraphicsEngine_drawGame.cpp:164,178,252 colour the points by
  (calculateState_J2, ConstitutiveRelation_calculateState_J2.cpp:4-26);
* the binning (MPM_F_CELL): the cell each point is binned into equals the
  truncation x / h of GridPoint_Mediator.h:35, and after every rebin the
  device order equals a CPU stable sort of the points by (brick, cell in
  brick) in their previous order followed by the stride-8 spread inside each
  brick -- bit for bit, in both precisions;
* grid nodes (mpm_grid_nodes, PhysicsEngine::getGridPoints and
  getGridPoints_Kernel, PhysicsEngine.h:65-66) against the oracle's grid
  phases (…_MP.cpp:41-265) on the plain path and its two-field grid and
  kernel grid (…_MP_Contact.cpp:121-518) on the contact path.
"""
import os
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

SPREAD = 8   # lanes between points of one cell in the binned order


@pytest.fixture(scope="module")
def mpm():
    import code_mpm_auxetics_amd as m
    m.lib()
    return m


def _j2(s):
    """1/2 dev:dev of Voigt 6-vectors (tensor shear), rows."""
    p = (s[:, 0] + s[:, 1] + s[:, 2]) / 3.0
    d = s[:, :3] - p[:, None]
    return 0.5 * (d ** 2).sum(axis=1) + (s[:, 3:] ** 2).sum(axis=1)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_j2_readback(mpm, prec):
    sc = mpm.scene.small_world(material="j2", speed=2.0, nx=2, ny=1)
    with mpm.PhysicsEngine(sc, precision=mpm.FP64 if prec == "fp64" else mpm.FP32) as eng:
        eng.run_steps(300)
        g = eng.getMaterialPoints(mpm._abi.MPM_F_STRESS | mpm._abi.MPM_F_EPSP | mpm._abi.MPM_F_J2)
    d = sc.dc == 0
    assert np.count_nonzero(g["eps_p"][d].any(axis=1)) > 0.1 * d.sum()   # yielded points
    for col, field in ((0, "stress"), (1, "eps_p")):
        want = _j2(g[field])
        assert np.abs(want[d]).max() > 0
        assert np.allclose(g["j2"][:, col], want, rtol=1e-12, atol=1e-15 * np.abs(want).max()), field


def _bin_keys(cell, S, nb):
    """Brick and cell-in-brick key of k_rebin_keys, as one int64."""
    b = (cell[:, 0] // S[0]) + nb[0] * ((cell[:, 1] // S[1]) + nb[1] * (cell[:, 2] // S[2]))
    lc = (cell[:, 0] % S[0]) + S[0] * ((cell[:, 1] % S[1]) + S[1] * (cell[:, 2] % S[2]))
    return b.astype(np.int64), (b.astype(np.int64) << 20) | lc


def _expected_order(prev_order, cell, S, nb):
    """Device position of every point after a rebin: stable sort of the
    points in their previous device order by (brick, cell in brick), then
    inside each brick of n points sorted position t = q*P + r moves to
    sum_{r' < r} ceil((n - r') / P) + q."""
    ids = np.argsort(prev_order, kind="stable")            # point ids in previous device order
    brick, key = _bin_keys(cell[ids], S, nb)
    srt = np.argsort(key, kind="stable")
    sorted_ids, sorted_brick = ids[srt], brick[srt]
    n = len(ids)
    out = np.empty(n, np.int64)
    starts = np.flatnonzero(np.r_[True, sorted_brick[1:] != sorted_brick[:-1]])
    ends = np.r_[starts[1:], n]
    for st, en in zip(starts, ends):
        cnt = en - st
        t = np.arange(cnt)
        r, q = t % SPREAD, t // SPREAD
        cum = np.cumsum([0] + [(cnt - rr - 1) // SPREAD + 1 if cnt > rr else 0 for rr in range(SPREAD)])
        out[sorted_ids[st:en]] = st + cum[r] + q
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_binning_cells_and_permutation_bitexact(mpm, prec):
    """A block in free flight through 40 x 40 x 20 cells rebins many times.
    Stepping one drift-check interval per call puts at most one rebin at the
    end of each call, so the order before it is the previous call's."""
    sc = mpm.scene.translation_world()
    K = 4
    F = mpm._abi
    h = np.array(sc.h)
    S = np.minimum([4, 4, 10], sc.cells)
    nb = np.array([(c + 1) // s + (1 if (c + 1) % s else 0) for c, s in zip(sc.cells, S)])
    rebins_checked = 0
    with mpm.PhysicsEngine(sc, precision=mpm.FP64 if prec == "fp64" else mpm.FP32, drift_check=K) as eng:
        g = eng.getMaterialPoints(F.MPM_F_X | F.MPM_F_CELL)
        prev_order = g["order"].astype(np.int64)
        assert np.array_equal(np.sort(prev_order), np.arange(sc.n))
        rebins = eng.checks()["rebins"]
        for _ in range(60):
            eng.run_steps(K)
            g = eng.getMaterialPoints(F.MPM_F_X | F.MPM_F_CELL)
            cell, order = g["cell"].astype(np.int64), g["order"].astype(np.int64)
            u = g["x"] / h
            tr = np.trunc(u).astype(np.int64)
            if prec == "fp64":
                assert np.array_equal(cell, tr)   # the same expression, GridPoint_Mediator.h:35
            else:
                # fixed-point cells; a read-back FP64 position within 1e-7 of a node
                # plane may round to the other side of it
                near = np.abs(u - np.round(u)) < 1e-7
                assert np.array_equal(cell[~near], tr[~near])
            c = eng.checks()
            if c["rebins"] > rebins:
                assert c["rebins"] == rebins + 1
                assert np.array_equal(order, _expected_order(prev_order, cell, S, nb))
                rebins_checked += 1
            else:
                assert np.array_equal(order, prev_order)
            rebins, prev_order = c["rebins"], order
    print(f"{prec}: {rebins_checked} rebins checked", flush=True)
    assert rebins_checked >= 3


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_grid_nodes_plain_vs_oracle(mpm, oracle_mod, prec):
    """getGridPoints after upload and after 50 steps: the grid of the next
    step (P2G of the current points, damping, boundary conditions, reaction,
    platen override) against the oracle's phases 1-6 from the same points."""
    sc = mpm.scene.small_world(material="j2", speed=2.0, nx=2, ny=1)
    fp64 = prec == "fp64"
    orc = oracle_mod.Oracle(sc, strain_form=0 if fp64 else 1)
    tol = 1e-12 if fp64 else 1e-4
    with mpm.PhysicsEngine(sc, precision=mpm.FP64 if fp64 else mpm.FP32) as eng:
        for steps in (0, 50):
            if steps:
                eng.run_steps(steps)
                orc.run_steps(steps)
            gg = eng.getGridPoints()
            og = orc.grid_state()
            act = og["mass"] > 0
            assert act.sum() > 0
            if fp64:
                assert np.array_equal(gg["mass"] > 0, act)
            for k in ("mass", "v", "f"):
                e = rel_l2(gg[k][act], og[k][act])
                print(f"{prec} step {steps}: {k} {e:.3e}", flush=True)
                # node forces are sums with cancellation: after steps, the
                # states' 1e-13 agreement shows up amplified in them
                assert e <= (tol if (k != "f" or steps == 0 or not fp64) else 1e-10), k
            fy = og["f_temp"][:, 1]
            assert np.abs(gg["f_temp"][:, 1] - fy).max() <= tol * max(np.abs(fy).max(), 1e-30)
            assert gg["mass"].sum() == pytest.approx(sc.mass[sc.dc == 0].sum(), rel=1e-6)
            assert np.array_equal(gg["fixed"], sc.fixed)
            assert np.allclose(gg["x"], mpm.scene.node_positions(sc.length, sc.cells), rtol=1e-15, atol=0)
    # the platen override: nodes under the top platen move with it
    top = gg["mass"] > 0
    assert np.any(np.abs(gg["v"][top, 1] + 2.0) < 1e-9)


def test_grid_nodes_contact_vs_oracle(mpm, oracle_mod):
    """After contact steps the nodes are that step's first field and the
    kernel grid, against the oracle's two-field step."""
    sc = make_golden.contact_blocks()
    orc = oracle_mod.Oracle(sc)
    orc.run_steps(2, 1)
    with mpm.PhysicsEngine(sc, precision=mpm.FP64) as eng:
        eng.run_steps(2, path=mpm.PATH_CONTACT)
        gg, gk = eng.getGridPoints(), eng.getGridPoints_Kernel()
    a = orc.a
    act = a["gm"] > 0
    assert act.sum() > 0
    for k, ok in (("mass", "gm"), ("v", "gv"), ("f", "gf")):
        e = rel_l2(gg[k][act], a[ok][act])
        print(f"contact {k} {e:.3e}", flush=True)
        assert e <= 1e-9, k
    assert rel_l2(gk["kernel"], a["gk"]) <= 1e-12
    assert rel_l2(gk["kernel_grad"], a["gkgrad"]) <= 1e-9
    assert np.abs(a["gkgrad"]).max() > 0
    # nodes the last step did not touch read as reset nodes (…_MP_Contact.cpp:41-67)
    idle = gg["active"] == 0
    assert idle.sum() > 0 and not act[idle].any()
    for k in ("mass", "v", "f"):
        assert not gg[k][idle].any(), k
