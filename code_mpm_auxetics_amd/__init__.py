"""MI355X-native explicit MPM step for auxetic-cell compression.

The GPU path lives in ``libmpm_auxetics.so`` (HIP, gfx950) behind the C ABI of
``include/mpm_auxetics.h``; ``PhysicsEngine`` mirrors the reference solver's
interface over that ABI, and ``scene`` builds the host-side point domains.
"""
from . import constitutive, dump, scene
from ._abi import MpmError, lib
from ._abi import MPM_F_ALL as F_ALL
from ._abi import MPM_PATH_CONTACT as PATH_CONTACT
from ._abi import MPM_PATH_PLAIN as PATH_PLAIN
from .engine import FP32, FP64, PhysicsEngine, sort_keys

__all__ = ["PhysicsEngine", "FP32", "FP64", "PATH_PLAIN", "PATH_CONTACT", "F_ALL", "MpmError", "scene", "dump",
           "constitutive", "sort_keys", "lib"]
