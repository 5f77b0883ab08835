"""Host mirror of the reference's ``ConstitutiveRelation`` over the GPU.

ConstitutiveRelation.h:16-34 declares one object per material point with
output members (``d6StressIncrement``, ``d6PlasticStrainIncrement``,
``dBackstress_IsotropicIncrement``, ``d_J2``) and one method per law.  Here
the same methods evaluate the step's own device functions through
``mpm_constitutive`` (FP64, the reference's evaluation order), and
``increments`` does the same for whole arrays of inputs at once.

Voigt order is xx, yy, zz, xy, yz, zx; strains carry engineering shear,
stresses tensor shear (…_MP.cpp:304-312).  F is a 3 x 3 matrix, or 9 values
column-major as glm stores it.  The ViscoElastic and IdealGas laws
(ConstitutiveRelation_calculateIncrement_ViscoElastic.cpp,
…_IdealGas.cpp) are never dispatched by the reference step
(…_MP.cpp:332-339); their methods keep the reference's signatures and raise.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import MpmCrBatch, check, lib

LAWS = {"elastic": _abi.MPM_LAW_ELASTIC, "neohookean": _abi.MPM_LAW_NEOHOOKEAN, "plastic": _abi.MPM_LAW_PLASTIC,
        "vonmises": _abi.MPM_LAW_VONMISES, "j2": _abi.MPM_LAW_J2_STATE,
        "viscoelastic": _abi.MPM_LAW_VISCOELASTIC, "idealgas": _abi.MPM_LAW_IDEALGAS}


def _arr(a, n, k):
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n, k) if k > 1 else (n,)))
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


NOALLOC_MAX = 8192   # inputs mpm_constitutive_noalloc takes per call


def increments(law, E=None, nu=None, stress=None, strain_inc=None, F=None, yield_stress=None, alpha=None,
               C0=None, C1=None, device=0, noalloc=False):
    """Apply one law to n inputs on the GPU.  ``law`` is a name of ``LAWS``
    or an ``MPM_LAW_*`` code; scalars broadcast.  Returns a dict with
    ``stress_inc`` (n, 6), ``plastic_inc`` (n, 6), ``alpha_inc`` (n,), or
    ``j2`` (n,) for the J2 state.  ``noalloc`` uses mpm_constitutive_noalloc
    (no device allocation, at most NOALLOC_MAX inputs; same results)."""
    code = LAWS[law] if isinstance(law, str) else int(law)
    sizes = [np.asarray(v).reshape(-1, k).shape[0] for v, k in ((stress, 6), (strain_inc, 6), (F, 9))
             if v is not None and np.asarray(v).ndim > 1]
    sizes += [np.asarray(v).size for v in (E, nu, yield_stress, alpha, C0, C1) if v is not None and np.ndim(v) > 0]
    n = max(sizes) if sizes else 1
    keep = []

    def inp(v, k):
        if v is None:
            return None
        a = _arr(np.asarray(v, dtype=np.float64).reshape(-1, k) if k > 1 else v, n, k)
        keep.append(a)
        return a

    b = MpmCrBatch()
    b.n = n
    b.E, b.nu, b.yield_ = _ptr(inp(E, 1)), _ptr(inp(nu, 1)), _ptr(inp(yield_stress, 1))
    b.alpha, b.C0, b.C1 = _ptr(inp(alpha, 1)), _ptr(inp(C0, 1)), _ptr(inp(C1, 1))
    b.stress, b.strain_inc = _ptr(inp(stress, 6)), _ptr(inp(strain_inc, 6))
    if F is not None:
        Fa = np.asarray(F, dtype=np.float64)
        if Fa.shape[-2:] == (3, 3):   # matrix [row][col] -> column-major 9
            Fa = np.swapaxes(Fa, -1, -2).reshape(-1, 9)
        b.F = _ptr(inp(Fa, 9))
    out = {}
    if code == _abi.MPM_LAW_J2_STATE:
        out["j2"] = np.zeros(n)
        b.j2 = _ptr(out["j2"])
    else:
        out["stress_inc"], out["plastic_inc"], out["alpha_inc"] = np.zeros((n, 6)), np.zeros((n, 6)), np.zeros(n)
        b.stress_inc, b.plastic_inc, b.alpha_inc = (_ptr(out["stress_inc"]), _ptr(out["plastic_inc"]),
                                                    _ptr(out["alpha_inc"]))
    entry = lib().mpm_constitutive_noalloc if noalloc else lib().mpm_constitutive
    check(entry(code, C.byref(b), device), None)
    return out


class ConstitutiveRelation:
    """ConstitutiveRelation.h:16-34: one call per point, results in the
    output members (ConstitutiveRelation.h:21-24).  Each call evaluates one
    input through mpm_constitutive_noalloc, so per-point use allocates no
    device memory."""

    def __init__(self, device=0):
        self.device = device
        self.d6StressIncrement = np.zeros(6)
        self.d6PlasticStrainIncrement = np.zeros(6)
        self.dBackstress_IsotropicIncrement = 0.0
        self.d_J2 = 0.0

    def _store(self, r):
        self.d6StressIncrement = r["stress_inc"][0].copy()
        self.d6PlasticStrainIncrement = r["plastic_inc"][0].copy()
        self.dBackstress_IsotropicIncrement = float(r["alpha_inc"][0])

    def calculateIncrement_Elastic(self, dE, dNu, d6StrainIncrement):
        self._store(increments("elastic", E=dE, nu=dNu, strain_inc=np.reshape(d6StrainIncrement, (1, 6)),
                               device=self.device, noalloc=True))

    def calculateIncrement_NeoHookean_6D(self, dE, dNu, d6Stress_Current, d33DeformationGradient):
        self._store(increments("neohookean", E=dE, nu=dNu, stress=np.reshape(d6Stress_Current, (1, 6)),
                               F=np.reshape(d33DeformationGradient, (1, 3, 3)), device=self.device, noalloc=True))

    def calculateIncrement_PerfectlyPlastic_6D(self, dE, dNu, dYield, d6StressCurrent, d6StrainIncrement):
        self._store(increments("plastic", E=dE, nu=dNu, yield_stress=dYield,
                               stress=np.reshape(d6StressCurrent, (1, 6)),
                               strain_inc=np.reshape(d6StrainIncrement, (1, 6)), device=self.device, noalloc=True))

    def calculateIncrement_VonMisesHardening_6D(self, dE, dNu, dYield, dBackstress_Isotropic,
                                                dHardening_Isotropic_C0, dHardening_Isotropic_C1,
                                                d6StressCurrent, d6StrainIncrement):
        self._store(increments("vonmises", E=dE, nu=dNu, yield_stress=dYield, alpha=dBackstress_Isotropic,
                               C0=dHardening_Isotropic_C0, C1=dHardening_Isotropic_C1,
                               stress=np.reshape(d6StressCurrent, (1, 6)),
                               strain_inc=np.reshape(d6StrainIncrement, (1, 6)), device=self.device, noalloc=True))

    def calculateState_J2(self, d6State):
        self.d_J2 = float(increments("j2", stress=np.reshape(d6State, (1, 6)), device=self.device,
                                     noalloc=True)["j2"][0])

    def calculateIncrement_ViscoElastic_6D(self, dE, dNu, dEta, d6Stress_Current, d6Strain_New, d6Strain_Rate_New):
        raise NotImplementedError("ViscoElastic is never dispatched by the reference step (…_MP.cpp:332-339)")

    def calculateIncrement_IdealGass(self, dHeatCapacityRatio, dSpecificHeat, dShearViscosity, dBulkViscosity,
                                     d6Stress_Current, dDensity, dTemperature, d6Strain_Rate):
        raise NotImplementedError("IdealGas is never dispatched by the reference step (…_MP.cpp:332-339)")
