"""Headless driver of the step: the role of Main.cpp and
GraphicsEngine::gameLoop (Main.cpp:14-27, GraphicsEngine_gameLoop.cpp:4-36)
without the window.

Builds a world (``scene.from_config``), reports its description
(initializeWorld_*'s reportConsole, …PolygonCell.cpp:275-312), then advances
frames of 1000 time increments (gameLoop's ``1.0e3 * getTime_Increment()``)
until the run reports completion.  Report lines go to stdout (``--verbose``)
and to the log file (``--log``, the reference's ./bmp/log.txt).  Every
``--snapshot-interval`` seconds of simulated time it writes
``Snapshot_<microseconds, 6 digits>.bmp`` (the gameLoop's naming): an
orthographic x-y view of the points colored as drawGame colors them
(GraphicsEngine_drawGame.cpp:58-68: platen white, surface blue, others red).

    python -m code_mpm_auxetics_amd.run --world reference_polygon --log bmp/log.txt --verbose
"""
import argparse
import os
import struct
import sys

import numpy as np

from . import _abi, scene
from .engine import FP32, FP64, PhysicsEngine

_RED, _BLUE, _WHITE = (255, 0, 0), (0, 0, 255), (255, 255, 255)


def write_bmp(path, rgb):
    """24-bit BMP of an (h, w, 3) uint8 image (row 0 at the top)."""
    h, w, _ = rgb.shape
    pad = (4 - (w * 3) % 4) % 4
    rows = np.ascontiguousarray(rgb[::-1, :, ::-1])   # bottom-up, BGR
    data = b"".join(r.tobytes() + b"\0" * pad for r in rows)
    with open(path, "wb") as f:
        f.write(b"BM" + struct.pack("<IHHI", 54 + len(data), 0, 0, 54))
        f.write(struct.pack("<IiiHHIIiiII", 40, w, h, 1, 24, 0, len(data), 2835, 2835, 0, 0))
        f.write(data)


def render(sc, x, vol, size=600):
    """Orthographic x-y view of the grid box: each point a square of
    0.8 * V^(1/3) (drawGame's particle size), drawn in point order."""
    L = sc.length
    scale = size / max(L[0], L[1])
    w, h = max(1, int(round(L[0] * scale))), max(1, int(round(L[1] * scale)))
    img = np.zeros((h, w, 3), dtype=np.uint8)
    col = np.tile(np.array(_RED, np.uint8), (sc.n, 1))
    col[sc.surface != 0] = _BLUE
    col[sc.dc != 0] = _WHITE
    half = np.maximum(0.4 * np.cbrt(np.abs(vol)) * scale, 0.5)
    cx, cy = x[:, 0] * scale, (L[1] - x[:, 1]) * scale
    x0 = np.clip(np.floor(cx - half), 0, w - 1).astype(int)
    x1 = np.clip(np.ceil(cx + half), 1, w).astype(int)
    y0 = np.clip(np.floor(cy - half), 0, h - 1).astype(int)
    y1 = np.clip(np.ceil(cy + half), 1, h).astype(int)
    for i in range(sc.n):
        img[y0[i]:y1[i], x0[i]:x1[i]] = col[i]
    return img


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--world", default="reference_polygon", help="world name or JSON spec file (scene.from_config)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--path", default="plain", choices=["plain", "contact"])
    ap.add_argument("--log", default=None, help="append report lines to this file (the reference: ./bmp/log.txt)")
    ap.add_argument("--verbose", action="store_true", help="print report lines")
    ap.add_argument("--snapshot-dir", default=None, help="write BMP snapshots here")
    ap.add_argument("--snapshot-interval", type=float, default=1.0e-4)
    ap.add_argument("--frames", type=int, default=0, help="stop after this many frames (0: run to time_end)")
    ap.add_argument("--frame-steps", type=float, default=1.0e3, help="time increments per frame")
    args = ap.parse_args(argv)

    sc = scene.from_config(args.world)
    eng = PhysicsEngine(sc, precision=FP64 if args.precision == "fp64" else FP32, verbose=args.verbose)
    if args.log:
        os.makedirs(os.path.dirname(os.path.abspath(args.log)), exist_ok=True)
        eng.set_log(args.log)
    eng.reportConsole(scene.describe(sc))
    if args.snapshot_dir:
        os.makedirs(args.snapshot_dir, exist_ok=True)
    step = (eng.runSimulation_Classic_SinglePass_MP if args.path == "plain"
            else eng.runSimulation_Classic_SinglePass_MP_Contact)
    t_last_snap, frames, done = 0.0, 0, 0
    while not done:
        t = eng.getTime_Current()
        if args.snapshot_dir and t - t_last_snap > args.snapshot_interval:
            t_last_snap = t
            pts = eng.getMaterialPoints(_abi.MPM_F_X | _abi.MPM_F_VOL)
            name = "Snapshot_%06d.bmp" % int(t * 1.0e6)
            write_bmp(os.path.join(args.snapshot_dir, name), render(sc, pts["x"], pts["vol"]))
        done = step(args.frame_steps * eng.getTime_Increment())
        frames += 1
        if args.frames and frames >= args.frames:
            break
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
