"""Committed golden fixtures (tests/golden/*.npz, written by make_golden.py):
the oracle must reproduce them (CPU), and the FP64 HIP path must match them
to the FP64 parity bound (GPU)."""
import os
import sys

import numpy as np
import pytest

from conftest import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

PLAIN = [k for k, v in make_golden.CASES.items() if v[1] == 0]


@pytest.mark.parametrize("name", list(make_golden.CASES))
def test_oracle_reproduces_fixture(oracle_mod, name):
    sc, path, steps, sf, out = make_golden.load(name)
    orc = oracle_mod.Oracle(sc, strain_form=sf)
    orc.run_steps(steps, path)
    p = orc.points()
    for k in make_golden.OUT_KEYS:
        ref = out[k]
        assert np.allclose(p[k], ref, rtol=1e-12, atol=1e-12 * (np.abs(ref).max() + 1e-300)), k
    assert len(orc.reaction) == len(out["reaction"])
    assert np.allclose(orc.reaction, out["reaction"], rtol=1e-10, atol=1e-10 * (np.abs(out["reaction"]).max() + 1e-300))


@pytest.mark.gpu
@pytest.mark.parametrize("name", PLAIN)
def test_fp64_gpu_matches_fixture(name):
    import code_mpm_auxetics_amd as m
    sc, path, steps, sf, out = make_golden.load(name)
    x0 = sc.x.copy()
    eng = m.PhysicsEngine(sc, precision=m.FP64, strain_form=sf)
    eng.run_steps(steps)
    g = eng.getMaterialPoints()
    ex = rel_l2(g["x"] - x0, out["x"] - x0)
    ev = rel_l2(g["v"], out["v"])
    es = rel_l2(g["stress"], out["stress"])
    print(f"{name}: rel L2 disp {ex:.3e} vel {ev:.3e} stress {es:.3e}", flush=True)
    assert ex <= 1e-9 and ev <= 1e-9 and es <= 1e-9
    # reaction (sum of fixed-y node forces): relative to its scale, with an
    # absolute floor for worlds where it is round-off only (|R| < 1e-20 N)
    hg, ho = eng.reaction_history(), out["reaction"]
    assert len(hg) == len(ho)
    assert np.abs(hg - ho).max() <= max(1e-8 * np.abs(ho).max(), 1e-20)
    eng.close()
