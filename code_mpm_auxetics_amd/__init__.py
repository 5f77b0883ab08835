"""MI355X-native explicit MPM step for auxetic-cell compression.

The GPU path lives in ``libmpm_auxetics.so`` (HIP, gfx950) behind the C ABI of
``include/mpm_auxetics.h``; ``PhysicsEngine`` mirrors the reference solver's
interface over that ABI, and ``scene`` builds the host-side point domains.
"""
from . import scene
from ._abi import MpmError, lib
from .engine import FP32, FP64, PhysicsEngine, sort_keys

__all__ = ["PhysicsEngine", "FP32", "FP64", "MpmError", "scene", "sort_keys", "lib"]
