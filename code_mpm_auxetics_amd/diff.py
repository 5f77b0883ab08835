"""Differentiable plain step: gradients of a loss on the final point
displacements with respect to the per-point Young's modulus and the initial
positions.

The forward and the adjoint run in the HIP library (``mpm_diff_*`` in
``include/mpm_auxetics.h``): FP64 steps of
PhysicsEngine_runSimulation_Classic_SinglePass_MP.cpp:41-368 (strain increment
L dt, damping, TimeLine platen) for elastic, NeoHookean (when the scene
dispatches it), perfectly plastic and hardening J2 points, with checkpoints
every ``checkpoint`` steps, and the hand-written adjoint of P2G, grid update,
G2P and the constitutive updates replayed from them.  The yield decision of
each point and step is held fixed, as the other discrete choices are.
The run carries the displacement u = x - x0 as state and returns it, so that
displacements many orders below the positions keep full FP64 precision.
``rollout`` exposes the pair as a ``torch.autograd.Function``; the effective
Poisson's ratio loss is ordinary torch code on top of it.

The reference has no gradients; the discrete choices of its step (the cell
of each point, mass gates, boundary masks, platen nodes) are held fixed, so
the gradient is that of the piecewise-smooth map the step defines.
"""
import ctypes as C

import numpy as np
import torch

from . import _abi
from ._abi import MpmConfig, MpmPoints, lib


def _ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def _check(code, h):
    if code != _abi.MPM_OK:
        msg = lib().mpm_diff_last_error(h)
        raise _abi.MpmError(code, msg.decode() if msg else "")


class DiffRun:
    """One differentiable run of ``n_steps`` steps of ``scene.dt_max`` on one
    GPU.  ``forward(E, x0)`` returns the final displacements (n, 3);
    ``backward(g)`` returns (dL/dE, dL/dx0) for dL/du_final = g of the last
    forward.  ``precision``: FP64 state, or FP32 state (u, v, F - I, sigma,
    alpha stored in float, positions x0 + u, every step's arithmetic and the
    adjoint in FP64, the rounding of the stored state taken as the identity)."""

    def __init__(self, scene, n_steps, checkpoint=20, device=0, precision=_abi.MPM_FP64):
        if precision not in (_abi.MPM_FP64, _abi.MPM_FP32):
            raise ValueError("precision must be FP64 or FP32")
        self.scene = scene
        self.precision = int(precision)
        self.n_steps = int(n_steps)
        self.checkpoint = int(checkpoint)
        cfg = MpmConfig()
        for a in range(3):
            cfg.length[a] = scene.length[a]
            cfg.cells[a] = scene.cells[a]
        cfg.dt_max = scene.dt_max
        cfg.time_start = 0.0
        cfg.time_end = scene.time_end
        cfg.damping = scene.damping
        cfg.mass_min = scene.mass_min
        cfg.precision = self.precision
        cfg.strain_form = 1
        cfg.neohookean = int(getattr(scene, "neohookean", 0))
        cfg.device = int(device)
        self._keep = {}
        tl_t = np.ascontiguousarray(scene.timeline_t, np.float64)
        tl_v = np.ascontiguousarray(scene.timeline_v, np.float64)
        self._keep.update(tl_t=tl_t, tl_v=tl_v)
        cfg.n_timeline = len(tl_t)
        cfg.timeline_t = _ptr(tl_t) if len(tl_t) else None
        cfg.timeline_v = _ptr(tl_v) if len(tl_t) else None
        fixed = np.ascontiguousarray(scene.fixed, np.uint8)
        self._keep["fixed"] = fixed
        cfg.node_fixed = _ptr(fixed, C.c_uint8)
        pts = MpmPoints()
        pts.n = scene.n
        for name, arr, ctype, dtype in (("x", scene.x, C.c_double, np.float64),
                                        ("v", scene.v, C.c_double, np.float64),
                                        ("fext", scene.fext, C.c_double, np.float64),
                                        ("mass", scene.mass, C.c_double, np.float64),
                                        ("vol0", scene.vol0, C.c_double, np.float64),
                                        ("vol", scene.vol0, C.c_double, np.float64),
                                        ("E", scene.E, C.c_double, np.float64),
                                        ("nu", scene.nu, C.c_double, np.float64),
                                        ("sy", scene.sy, C.c_double, np.float64),
                                        ("C0", scene.C0, C.c_double, np.float64),
                                        ("C1", scene.C1, C.c_double, np.float64),
                                        ("type", scene.type, C.c_int32, np.int32),
                                        ("dc", scene.dc, C.c_uint8, np.uint8)):
            a = np.ascontiguousarray(arr, dtype)
            self._keep["p_" + name] = a
            setattr(pts, name, _ptr(a, ctype))
        h = C.c_void_p()
        code = lib().mpm_diff_create(C.byref(cfg), C.byref(pts), self.n_steps, self.checkpoint, C.byref(h))
        if code != _abi.MPM_OK:
            raise _abi.MpmError(code, "mpm_diff_create failed (invalid scene or no device)")
        self._h = h
        self.n = scene.n

    def forward(self, E, x0):
        E = np.ascontiguousarray(E, np.float64).reshape(self.n)
        x0 = np.ascontiguousarray(x0, np.float64).reshape(self.n, 3)
        out = np.empty((self.n, 3))
        _check(lib().mpm_diff_forward(self._h, _ptr(E), _ptr(x0), _ptr(out)), self._h)
        return out

    def backward(self, g_ufinal):
        g = np.ascontiguousarray(g_ufinal, np.float64).reshape(self.n, 3)
        gE = np.empty(self.n)
        gx = np.empty((self.n, 3))
        _check(lib().mpm_diff_backward(self._h, _ptr(g), _ptr(gE), _ptr(gx)), self._h)
        return gE, gx

    def close(self):
        if getattr(self, "_h", None):
            lib().mpm_diff_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class _Rollout(torch.autograd.Function):
    """u_final = rollout(E, x0) with the library's adjoint as backward."""

    @staticmethod
    def forward(ctx, E, x0, run):
        ctx.run = run
        ctx.dev = E.device
        u = run.forward(E.detach().cpu().double().numpy(), x0.detach().cpu().double().numpy())
        return torch.from_numpy(u).to(device=E.device)

    @staticmethod
    def backward(ctx, g):
        g = g.detach().cpu().double()
        gE, gx = ctx.run.backward(g.numpy())
        return torch.from_numpy(gE).to(ctx.dev), torch.from_numpy(gx).to(ctx.dev), None


def rollout(run, E, x0):
    """Final displacements (n, 3) of ``run`` from moduli ``E`` (n,) and
    initial positions ``x0`` (n, 3), differentiable in both (platen rows are
    the platen's displacement, with zero gradient)."""
    return _Rollout.apply(E, x0, run)


def boundary_sets(scene, band_cells=1.0):
    """Point index sets of the loaded body's boundaries from the initial
    positions: bottom and top bands (axial), and left and right bands within
    the middle half of the height (lateral)."""
    keep = scene.dc == 0
    x = scene.x
    band = band_cells * scene.h[0]
    xs, ys = x[:, 0], x[:, 1]
    xmin, xmax = xs[keep].min(), xs[keep].max()
    ymin, ymax = ys[keep].min(), ys[keep].max()
    hgt = ymax - ymin
    mid = keep & (ys >= ymin + 0.25 * hgt) & (ys <= ymin + 0.75 * hgt)
    return dict(bottom=np.flatnonzero(keep & (ys <= ymin + band)), top=np.flatnonzero(keep & (ys >= ymax - band)),
                left=np.flatnonzero(mid & (xs <= xmin + band)), right=np.flatnonzero(mid & (xs >= xmax - band)))


def effective_poisson(x0, u, sets):
    """-(lateral strain) / (axial strain): each strain is the difference of
    the mean displacements ``u`` of two opposite boundary sets over their
    initial distance (torch tensors or numpy arrays)."""
    def strain(a, lo, hi):
        gap0 = x0[sets[hi], a].mean() - x0[sets[lo], a].mean()
        return (u[sets[hi], a].mean() - u[sets[lo], a].mean()) / gap0
    return -strain(0, "left", "right") / strain(1, "bottom", "top")


def poisson_loss(x0, u, sets, target=-0.6):
    """(nu_eff - target)^2."""
    return (effective_poisson(x0, u, sets) - target) ** 2
