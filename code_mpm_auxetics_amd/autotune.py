"""Autotuner for the launch shape of the step: brick size (cells per brick,
which sets the LDS arena and node window of the fused step kernel) and the
batch length (steps between drift checks, one captured hipGraph per batch).

Each candidate is timed on the caller's own scene with the library's device
timer (HIP events around the batch; ``getTime_Runtime``) after a warm-up, and
the fastest is returned.  Candidates the library rejects (an arena that does
not fit the LDS) are skipped.  Brick shape and batch length do not change the
physics; results differ only by the rounding of the fixed-point P2G sums.
"""
import json
import os

from . import _abi
from .engine import FP32, PhysicsEngine

# cells per brick (x, y, z); z is clamped to the grid by the library
DEFAULT_BRICKS = [(4, 4, 10), (4, 4, 5), (3, 4, 10), (4, 3, 10), (3, 3, 10), (4, 4, 8), (5, 4, 10), (5, 5, 10)]
DEFAULT_BATCHES = [8, 32]


def tune(scene, precision=FP32, bricks=None, batches=None, steps=32, warmup=8, repeats=2, verbose=False):
    """Time every (brick, batch) candidate on ``scene``; returns a dict with
    ``brick``, ``drift_check``, ``ms_per_step`` of the fastest and the table
    of all candidates (ms per step, or the error that ruled it out)."""
    bricks = bricks or DEFAULT_BRICKS
    batches = batches or DEFAULT_BATCHES
    table = []
    best = None
    for brick in bricks:
        for batch in batches:
            entry = {"brick": list(brick), "drift_check": batch}
            try:
                eng = PhysicsEngine(scene, precision=precision, brick=brick, drift_check=batch, track_strain=False)
            except _abi.MpmError as e:
                entry["error"] = str(e)
                table.append(entry)
                continue
            try:
                eng.run_steps(warmup)
                times = []
                for _ in range(repeats):
                    eng.run_steps(steps)
                    times.append(eng.getTime_Runtime() * 1e3 / steps)
                entry["ms_per_step"] = min(times)
            except _abi.MpmError as e:
                entry["error"] = str(e)
            finally:
                eng.close()
            table.append(entry)
            if verbose:
                print(json.dumps(entry), flush=True)
            if "ms_per_step" in entry and (best is None or entry["ms_per_step"] < best["ms_per_step"]):
                best = entry
    if best is None:
        raise RuntimeError("no candidate launch shape ran: " + json.dumps(table))
    return {"brick": tuple(best["brick"]), "drift_check": best["drift_check"], "ms_per_step": best["ms_per_step"],
            "table": table}


def save(result, path):
    """Write a tuning result as JSON (the table included)."""
    d = dict(result)
    d["brick"] = list(d["brick"])
    with open(path, "w") as f:
        json.dump(d, f, indent=1)


def load(path):
    """(brick, drift_check) of a saved tuning result, or None if absent."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return tuple(d["brick"]), int(d["drift_check"])
