"""mpm_constitutive_noalloc (the batch staged in the library's fixed device
array, no allocation) against mpm_constitutive (a staging buffer the library
allocates): bit-identical outputs on two calls in a row, the second with a
smaller batch and another law, so that the array's leftover contents from the
first call cannot leak into the second; batches beyond the array's 8192
inputs are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import cr_inputs as ci

pytestmark = pytest.mark.gpu


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _run(entry, law, n, E=None, nu=None, sy=None, alpha=None, C0=None, C1=None, s=None, de=None, F=None):
    from code_mpm_auxetics_amd import _abi as A
    b = A.MpmCrBatch()
    b.n = n
    keep = [np.ascontiguousarray(x, np.float64) if x is not None else None for x in (E, nu, sy, alpha, C0, C1, s, de, F)]
    b.E, b.nu, b.yield_, b.alpha, b.C0, b.C1, b.stress, b.strain_inc, b.F = [_dp(x) for x in keep]
    out = {}
    if law == A.MPM_LAW_J2_STATE:
        out["j2"] = np.full(n, np.nan)
        b.j2 = _dp(out["j2"])
    else:
        out["ds"], out["dep"], out["da"] = np.full((n, 6), np.nan), np.full((n, 6), np.nan), np.full(n, np.nan)
        b.stress_inc, b.plastic_inc, b.alpha_inc = _dp(out["ds"]), _dp(out["dep"]), _dp(out["da"])
    rc = entry(law, C.byref(b), 0)
    return rc, out


def test_noalloc_matches_allocating_entry_point_on_two_calls():
    import code_mpm_auxetics_amd as m
    from code_mpm_auxetics_amd import _abi as A
    L = m.lib()
    print("call 1: J2 hardening, 6000 inputs", flush=True)
    E, nu, sy, s, de = ci.inputs(4, n=6000)
    alpha, C0, C1 = ci.hardening(5, sy, n=6000)
    args = dict(E=E, nu=nu, sy=sy, alpha=alpha, C0=C0, C1=C1, s=s, de=de)
    rc0, ref = _run(L.mpm_constitutive, A.MPM_LAW_VONMISES, 6000, **args)
    rc1, got = _run(L.mpm_constitutive_noalloc, A.MPM_LAW_VONMISES, 6000, **args)
    assert rc0 == A.MPM_OK and rc1 == A.MPM_OK, L.mpm_last_error(None)
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k
    assert np.count_nonzero(ref["da"]) > 600                 # both yield branches ran
    print("call 2: NeoHookean, 700 inputs", flush=True)
    E2, nu2, s2, Fm = ci.neohookean_inputs(6, n=700)
    F2 = np.swapaxes(Fm, -1, -2).reshape(-1, 9)
    rc0, ref = _run(L.mpm_constitutive, A.MPM_LAW_NEOHOOKEAN, 700, E=E2, nu=nu2, s=s2, F=F2)
    rc1, got = _run(L.mpm_constitutive_noalloc, A.MPM_LAW_NEOHOOKEAN, 700, E=E2, nu=nu2, s=s2, F=F2)
    assert rc0 == A.MPM_OK and rc1 == A.MPM_OK, L.mpm_last_error(None)
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k
    print("call 3: J2 state, 300 inputs", flush=True)
    rc0, ref = _run(L.mpm_constitutive, A.MPM_LAW_J2_STATE, 300, s=s[:300])
    rc1, got = _run(L.mpm_constitutive_noalloc, A.MPM_LAW_J2_STATE, 300, s=s[:300])
    assert rc0 == A.MPM_OK and rc1 == A.MPM_OK
    assert np.array_equal(got["j2"], ref["j2"])


def test_noalloc_refuses_batches_beyond_its_array():
    import code_mpm_auxetics_amd as m
    from code_mpm_auxetics_amd import _abi as A
    L = m.lib()
    E, nu, _, _, de = ci.inputs(1, n=8193)
    rc, out = _run(L.mpm_constitutive_noalloc, A.MPM_LAW_ELASTIC, 8193, E=E, nu=nu, de=de)
    assert rc == A.MPM_E_ARG
    assert b"at most 8192" in L.mpm_last_error(None)
    assert np.isnan(out["ds"]).all()                          # nothing was written
    rc, out = _run(L.mpm_constitutive_noalloc, A.MPM_LAW_ELASTIC, 8192, E=E[:8192], nu=nu[:8192], de=de[:8192])
    assert rc == A.MPM_OK and np.isfinite(out["ds"]).all()
