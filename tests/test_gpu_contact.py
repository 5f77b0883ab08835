"""GPU contact path (runSimulation_Classic_SinglePass_MP_Contact) against the
FP64 oracle's two-field step: FP64 to the FP64 parity bound, FP32 to the
FP32 bound, plus determinism, the point kernel readback and switching paths
within one run."""
import os
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402


@pytest.fixture(scope="module")
def mpm():
    import code_mpm_auxetics_amd as m
    m.lib()
    return m


def _cmp(g, o, x0, tol, sel=slice(None)):
    ex = rel_l2((g["x"] - x0)[sel], (o["x"] - x0)[sel])
    ev = rel_l2(g["v"][sel], o["v"][sel])
    es = rel_l2(g["stress"][sel], o["stress"][sel])
    print(f"rel L2 disp {ex:.3e} vel {ev:.3e} stress {es:.3e}", flush=True)
    assert ex <= tol and ev <= tol and es <= tol, (ex, ev, es)


def test_contact_fixture_fp64(mpm):
    sc, path, steps, sf, out = make_golden.load("contact_blocks")
    x0 = sc.x.copy()
    eng = mpm.PhysicsEngine(sc, precision=mpm.FP64)
    eng.run_steps(steps, path=mpm.PATH_CONTACT)
    _cmp(eng.getMaterialPoints(), out, x0, 1e-9)
    eng.close()


def test_contact_blocks_first_steps_fp64(mpm, oracle_mod):
    """The steps in which contact engages match the oracle to the FP64 bound,
    including the point kernel (d_Kernel, d3_Kernel_Gradient).  Once a node's
    two fields have been corrected, their normal velocities are equal and the
    reference's approach test (…_MP_Contact.cpp:455) compares round-off, so
    later steps are checked by invariants (next test)."""
    sc = make_golden.contact_blocks()
    x0 = sc.x.copy()
    eng = mpm.PhysicsEngine(sc, precision=mpm.FP64)
    orc = oracle_mod.Oracle(sc)
    eng.run_steps(2, path=mpm.PATH_CONTACT)
    orc.run_steps(2, 1)
    g = eng.getMaterialPoints(mpm.F_ALL)
    o = orc.points()
    _cmp(g, o, x0, 1e-9)
    assert rel_l2(g["kernel"], orc.a["kernel"]) <= 1e-12
    assert rel_l2(g["kernel_grad"], orc.a["kernel_grad"]) <= 1e-9
    eng.close()


def test_contact_blocks_invariants(mpm, oracle_mod):
    """40 contact steps of the closing blocks: total momentum is conserved
    (equal and opposite contact forces, no boundary conditions), the state
    stays finite, and the aggregate motion of the facing layers (their
    closing speed) follows the oracle's to 1e-5 although individual contact
    decisions may follow round-off differently."""
    sc = make_golden.contact_blocks()
    p0 = (sc.mass[:, None] * sc.v).sum(axis=0)
    scale = np.abs(sc.mass[:, None] * sc.v).sum()
    eng = mpm.PhysicsEngine(sc, precision=mpm.FP64)
    eng.run_steps(40, path=mpm.PATH_CONTACT)
    vc = eng.getMaterialPoints()["v"]
    eng.close()
    orc = oracle_mod.Oracle(sc)
    orc.run_steps(40, 1)
    vo = orc.points()["v"]
    assert np.isfinite(vc).all()
    p1 = (sc.mass[:, None] * vc).sum(axis=0)
    assert np.allclose(p1, p0, rtol=0, atol=1e-10 * scale)
    nA = len(sc.x) // 2
    fa = (sc.surface == 1) & (np.arange(len(sc.x)) < nA)
    fb = (sc.surface == 1) & (np.arange(len(sc.x)) >= nA)
    closing_g = vc[fa, 0].mean() - vc[fb, 0].mean()
    closing_o = vo[fa, 0].mean() - vo[fb, 0].mean()
    print(f"closing speed gpu {closing_g:.9f} oracle {closing_o:.9f}", flush=True)
    assert abs(closing_g - closing_o) <= 1e-5 * abs(closing_o)
    assert rel_l2(vc, vo) <= 1e-3


@pytest.mark.parametrize("material", ["elastic", "j2"])
def test_contact_world_with_platens_fp64(mpm, oracle_mod, material):
    """Platen points take part in the contact P2G and override both fields;
    the reaction history is the platen's (…_MP_Contact.cpp:494-518)."""
    sc = mpm.scene.small_world(material=material, speed=(5.0 if material == "elastic" else 2.0))
    # the contact step moves platens with their stored velocity (no TimeLine,
    # …_MP_Contact.cpp:638-644): the upper platen closes at 2 m/s
    top = (sc.dc == 1) & (sc.x[:, 1] > sc.x[:, 1].mean())
    sc.v[top] = (0.0, -2.0, 0.0)
    x0 = sc.x.copy()
    orc = oracle_mod.Oracle(sc)
    orc.run_steps(60, 1)
    with mpm.PhysicsEngine(sc, precision=mpm.FP64) as eng:
        eng.run_steps(60, path=mpm.PATH_CONTACT)
        pts, hg = eng.getMaterialPoints(), eng.reaction_history()
    _cmp(pts, orc.points(), x0, 1e-9)
    ho = orc.reaction
    assert len(hg) == len(ho)
    assert np.abs(hg - ho).max() <= 1e-8 * np.abs(ho).max() + 1e-20
    assert np.abs(ho).max() > 0


def test_contact_fp32_bound(mpm, oracle_mod):
    """FP32 state on the contact path (platens in the P2G, no surface
    points) against the FP64 oracle's contact step, L dt strain form."""
    sc = mpm.scene.small_world(material="j2", speed=2.0)
    top = (sc.dc == 1) & (sc.x[:, 1] > sc.x[:, 1].mean())
    sc.v[top] = (0.0, -2.0, 0.0)
    x0 = sc.x.copy()
    eng = mpm.PhysicsEngine(sc, precision=mpm.FP32)
    orc = oracle_mod.Oracle(sc, strain_form=1)
    eng.run_steps(100, path=mpm.PATH_CONTACT)
    orc.run_steps(100, 1)
    _cmp(eng.getMaterialPoints(), orc.points(), x0, 1e-4, sel=sc.dc == 0)
    eng.close()


def test_contact_deterministic(mpm):
    sc = make_golden.contact_blocks()
    outs = []
    for _ in range(2):
        eng = mpm.PhysicsEngine(sc, precision=mpm.FP64)
        eng.run_steps(20, path=mpm.PATH_CONTACT)
        outs.append(eng.getMaterialPoints())
        eng.close()
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_switching_paths(mpm, oracle_mod):
    """plain, contact, plain within one run matches the oracle doing the same."""
    sc = mpm.scene.small_world(material="j2", speed=2.0)
    x0 = sc.x.copy()
    eng = mpm.PhysicsEngine(sc, precision=mpm.FP64)
    orc = oracle_mod.Oracle(sc)
    for path, n in ((0, 20), (1, 20), (0, 20)):
        eng.run_steps(n, path=path)
        orc.run_steps(n, path)
    _cmp(eng.getMaterialPoints(), orc.points(), x0, 1e-9)
    eng.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("world", ["blocks", "platens"])
def test_carried_bounds_equal_recomputed(mpm, prec, world):
    """The fixed-point scales of a contact step come from bounds over the
    points' contributions.  Within one call the previous step's G2P forms
    the deformable points' bounds from the state it writes; the first step
    of every call forms them in the kernel-density pass from the stored
    state.  Both must give the same bits: 25 steps in one call equal 25
    calls of one step, field by field."""
    if world == "blocks":
        sc = make_golden.contact_blocks()
    else:
        sc = mpm.scene.small_world(material="j2", speed=2.0)
    # a power-of-two step below both worlds' limits: 25 of them add up
    # exactly, so the one call and the 25 calls take the same 25 steps
    sc.dt_max = 2.0 ** -26
    p = mpm.FP64 if prec == "fp64" else mpm.FP32
    out = []
    for calls in (1, 25):
        with mpm.PhysicsEngine(sc, precision=p) as eng:
            for _ in range(calls):
                eng.run_steps(25 // calls, path=mpm.PATH_CONTACT)
            out.append(eng.getMaterialPoints())
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def _thick_blocks(jitter=0.15, seed=5):
    """Two jittered blocks of 5 x 10 x 10 cells (2 points per cell and axis)
    closing across half a cell, facing layers flagged as surface, from the
    grid's origin, so that a brick of 8 x 8 x 8 cells holds points in all of
    its cells."""
    h, cells, off = 0.001, (16, 12, 12), 0.0005

    def block(x0, nx):
        xs = [(x0 + i) * h + (k + 0.5) * off for i in range(nx) for k in range(2)]
        ys = [i * h + (k + 0.5) * off for i in range(10) for k in range(2)]
        zs = [i * h + (k + 0.5) * off for i in range(10) for k in range(2)]
        return np.array([(a, b, c) for a in xs for b in ys for c in zs])
    A, B = block(0, 5), block(5.5, 5)
    X = np.vstack([A, B])
    n = len(X)
    vel = np.zeros((n, 3))
    vel[: len(A), 0] = 2.0
    vel[len(A):, 0] = -2.0
    surf = np.zeros(n, np.uint8)
    surf[: len(A)][A[:, 0] > A[:, 0].max() - 0.6 * off] = 1
    surf[len(A):][B[:, 0] < B[:, 0].min() + 0.6 * off] = 1
    X = X + np.random.default_rng(seed).uniform(-jitter, jitter, X.shape) * off
    vol = off ** 3
    return mpm_scene().Scene(length=tuple(c * h for c in cells), cells=cells,
                             fixed=np.zeros(np.prod([c + 1 for c in cells]), np.uint8), x=X,
                             mass=np.full(n, 1000 * vol), vol0=np.full(n, vol), E=np.full(n, 1e7),
                             nu=np.full(n, 0.3), sy=np.zeros(n), C0=np.zeros(n), C1=np.zeros(n),
                             type=np.zeros(n, np.int32), dc=np.zeros(n, np.uint8), v=vel, surface=surf,
                             dt_max=1e-7, time_end=1.0, damping=0.0, mass_min=0.0)


def mpm_scene():
    from code_mpm_auxetics_amd import scene
    return scene


def test_large_bricks_take_the_global_sums(mpm):
    """With bricks of 8 x 8 x 8 cells the point passes' node boxes (up to
    9 x 9 x 9 = 729 nodes, 11^3 = 1331 for the kernel density) exceed their
    LDS windows (416 and 512 nodes, 1024), so those blocks add their terms
    with global atomics.  The contact path's sums are fixed-point integers
    with step-wide scales and its choices are order-free, so in FP64 state
    the result equals the windowed run's (bricks of 4 x 4 x 4) bit for bit."""
    sc = _thick_blocks()
    out = []
    for brick in ((4, 4, 4), (8, 8, 8)):
        with mpm.PhysicsEngine(sc, precision=mpm.FP64, brick=brick) as eng:
            eng.run_steps(20, path=mpm.PATH_CONTACT)
            out.append(eng.getMaterialPoints(mpm.F_ALL))
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def _sliding_blocks():
    """The closing blocks with a tangential relative velocity (A +y, B -y)."""
    sc = make_golden.contact_blocks()
    nA = len(sc.x) // 2
    sc.v[:nA, 1] = 1.0
    sc.v[nA:, 1] = -1.0
    return sc, nA


def test_friction_zero_is_the_reference_path(mpm):
    sc, _ = _sliding_blocks()
    out = []
    for kw in ({}, {"friction": 0.0}):
        eng = mpm.PhysicsEngine(sc, precision=mpm.FP64, **kw)
        eng.run_steps(20, path=mpm.PATH_CONTACT)
        out.append(eng.getMaterialPoints())
        eng.close()
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_friction_conserves_momentum_and_adds_tangential_coupling(mpm):
    """Coulomb friction (an extension; the reference's correction is normal
    only): the capped tangential impulse is applied equal and opposite to the
    two fields, so total momentum stays conserved (no boundary conditions,
    no mass gate in this scene), and the tangential momentum block A gives up
    to block B grows with the coefficient.  Most of the blocks' coupling goes
    through nodes both blocks share in one field, which the reference treats
    as no-slip, so the frictional share is a fraction of the total."""
    sc, nA = _sliding_blocks()
    m = sc.mass
    p0 = (m[:, None] * sc.v).sum(axis=0)
    scale = np.abs(m[:, None] * sc.v).sum()
    pa0 = (m[:nA, None] * sc.v[:nA]).sum(axis=0)
    dy = {}
    for mu in (0.0, 0.2, 5.0):
        eng = mpm.PhysicsEngine(sc, precision=mpm.FP64, friction=mu)
        eng.run_steps(40, path=mpm.PATH_CONTACT)
        v = eng.getMaterialPoints()["v"]
        eng.close()
        assert np.isfinite(v).all()
        p1 = (m[:, None] * v).sum(axis=0)
        assert np.allclose(p1, p0, rtol=0, atol=1e-12 * scale), (mu, p1 - p0)
        da = (m[:nA, None] * v[:nA]).sum(axis=0) - pa0
        assert da[0] < 0.0            # the normal impulse slows A
        dy[mu] = -da[1]               # tangential momentum A loses (it slides in +y)
        print(f"mu {mu}: A normal {da[0]:+.4e} tangential {da[1]:+.4e}", flush=True)
    assert dy[0.0] > 0.0 and dy[5.0] > 1.03 * dy[0.0] and dy[0.2] >= dy[0.0]
