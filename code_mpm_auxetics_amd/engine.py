"""Python mirror of the reference ``PhysicsEngine`` interface over the C ABI.

Method names follow PhysicsEngine.h:45-66 so that a driver written against the
reference reads the same:

* ``runSimulation_Classic_SinglePass_MP(dt_total)`` -> 0 / 1  (PhysicsEngine.h:45)
* ``runSimulation_Classic_SinglePass_MP_Contact(dt_total)``    (PhysicsEngine.h:46)
* ``getMaterialPoints()`` -> dict of numpy arrays, original order (PhysicsEngine.h:64)
* ``getTime_Current / getTime_End / getTime_Increment / getTime_Runtime`` (:57-60)
* ``getCount_MaterialPoint / getCount_GridPoint`` (:62-63)
* ``reportConsole(description)`` -> the console/log line (PhysicsEngine.cpp:33-120)
* ``getScript_MaterialPoints / getScript_GridPoints`` -> the factories' text dumps
  (MaterialPoint_Factory_getScript.cpp:4-28, GridPoint_Factory.h:69-100)

All compute runs in the HIP library; this module only marshals arrays.
"""
import ctypes as C

import numpy as np

from . import _abi, dump
from ._abi import MpmConfig, MpmNodes, MpmPoints, check, lib

FP64 = _abi.MPM_FP64
FP32 = _abi.MPM_FP32


def _ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def _c(a, dtype=np.float64):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


class PhysicsEngine:
    """One solver handle on one GPU (one HIP stream)."""

    def __init__(self, scene, precision=FP64, device=0, brick=(0, 0, 0), drift_check=8,
                 track_strain=True, strain_form=0, verbose=False, friction=0.0, profile=False,
                 graph=True, persistent=False):
        self.scene = scene
        L = lib()
        cfg = MpmConfig()
        for a in range(3):
            cfg.length[a] = scene.length[a]
            cfg.cells[a] = scene.cells[a]
            cfg.brick[a] = brick[a]
        cfg.dt_max = scene.dt_max
        cfg.time_start = 0.0
        cfg.time_end = scene.time_end
        cfg.damping = scene.damping
        cfg.mass_min = scene.mass_min
        cfg.console_interval = scene.console_interval
        cfg.precision = int(precision)
        cfg.strain_form = int(strain_form)
        cfg.neohookean = int(scene.neohookean)
        cfg.track_strain = 1 if track_strain else 0
        self._tl_t = _c(scene.timeline_t)
        self._tl_v = _c(scene.timeline_v)
        cfg.n_timeline = len(self._tl_t)
        cfg.timeline_t = _ptr(self._tl_t) if len(self._tl_t) else None
        cfg.timeline_v = _ptr(self._tl_v) if len(self._tl_t) else None
        self._fixed = _c(scene.fixed, np.uint8)
        cfg.node_fixed = _ptr(self._fixed, C.c_uint8)
        cfg.drift_check = int(drift_check)
        cfg.friction = float(friction)
        cfg.device = int(device)
        cfg.verbose = 1 if verbose else 0
        cfg.profile = 1 if profile else 0
        cfg.launch = _abi.MPM_LAUNCH_GRAPH if graph else _abi.MPM_LAUNCH_DIRECT
        cfg.persistent = 1 if persistent else 0
        if scene.node_group is not None and scene.n_groups > 0:
            self._group = np.ascontiguousarray(scene.node_group, np.uint16)
            cfg.node_group = self._group.ctypes.data_as(C.POINTER(C.c_uint16))
            cfg.n_groups = int(scene.n_groups)
        self.precision = precision
        self.track_strain = track_strain
        h = C.c_void_p()
        check(L.mpm_create(C.byref(cfg), C.byref(h)), None)
        self._h = h
        self._upload()

    # ------------------------------------------------------------------ setup
    def _upload(self):
        s = self.scene
        n = s.n
        keep = {}
        pts = MpmPoints()
        pts.n = n

        def put(name, arr, ctype=C.c_double, dtype=np.float64):
            a = _c(arr, dtype)
            keep[name] = a
            setattr(pts, name, _ptr(a, ctype))

        put("x", s.x)
        put("v", s.v)
        put("fext", s.fext)
        put("mass", s.mass)
        put("vol0", s.vol0)
        put("vol", s.vol0)
        for k in ("E", "nu", "sy", "C0", "C1"):
            put(k, getattr(s, k))
        put("type", s.type, C.c_int32, np.int32)
        put("dc", s.dc, C.c_uint8, np.uint8)
        put("surface", s.surface, C.c_uint8, np.uint8)
        check(lib().mpm_upload_points(self._h, C.byref(pts)), self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().mpm_destroy(self._h)
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

    # ------------------------------------------------------------------ steps
    def runSimulation_Classic_SinglePass_MP(self, dTimeIncrement_Total):
        fin = C.c_int32(0)
        check(lib().mpm_step(self._h, float(dTimeIncrement_Total), _abi.MPM_PATH_PLAIN, C.byref(fin)), self._h)
        return int(fin.value)

    def runSimulation_Classic_SinglePass_MP_Contact(self, dTimeIncrement_Total):
        fin = C.c_int32(0)
        check(lib().mpm_step(self._h, float(dTimeIncrement_Total), _abi.MPM_PATH_CONTACT, C.byref(fin)), self._h)
        return int(fin.value)

    def step_async(self, dTimeIncrement_Total, path=_abi.MPM_PATH_PLAIN):
        """Start the same advance as runSimulation_Classic_SinglePass_MP(_Contact)
        on the handle's worker thread and return at once; collect it with
        step_wait().  Other calls on the engine raise (MPM_E_STATE) meanwhile."""
        check(lib().mpm_step_async(self._h, float(dTimeIncrement_Total), path), self._h)

    def step_poll(self):
        """True once the asynchronous advance has finished."""
        done = C.c_int32(0)
        check(lib().mpm_step_poll(self._h, C.byref(done)), self._h)
        return bool(done.value)

    def step_wait(self):
        """Wait for the asynchronous advance; returns what the blocking call
        would have returned (1 when d_Time >= d_TimeEnd, else 0)."""
        fin = C.c_int32(0)
        check(lib().mpm_step_wait(self._h, C.byref(fin)), self._h)
        return int(fin.value)

    def run_steps(self, nsteps, path=_abi.MPM_PATH_PLAIN):
        """One call advancing dt_total = nsteps * dt_max, sub-stepped as the
        reference loop does (…_MP.cpp:18-27)."""
        fin = C.c_int32(0)
        check(lib().mpm_step(self._h, nsteps * self.scene.dt_max, path, C.byref(fin)), self._h)
        return int(fin.value)

    # ------------------------------------------------------------- readback
    def getMaterialPoints(self, mask=_abi.MPM_F_ALL):
        """The MaterialPoint fields of ``mask`` in the original point order.
        With MPM_F_J2: ``j2`` (n, 2), calculateState_J2 of the stress and of
        the plastic strain; with MPM_F_CELL: ``cell`` (n, 3) the binning cell
        and ``order`` (n,) the position in the binned device order."""
        n = self.scene.n
        if not self.track_strain:
            mask &= ~_abi.MPM_F_STRAIN
        out = dict(x=np.zeros((n, 3)), v=np.zeros((n, 3)), F=np.zeros((n, 9)), stress=np.zeros((n, 6)),
                   strain=np.zeros((n, 6)), eps_p=np.zeros((n, 6)), alpha=np.zeros(n), vol=np.zeros(n))
        if mask & _abi.MPM_F_KERNEL:
            out["kernel"] = np.zeros(n)
            out["kernel_grad"] = np.zeros((n, 3))
        if mask & _abi.MPM_F_J2:
            out["j2"] = np.zeros((n, 2))
        pts = MpmPoints()
        pts.n = n
        for k, a in out.items():
            setattr(pts, k, _ptr(a))
        if mask & _abi.MPM_F_CELL:
            out["cell"] = np.zeros((n, 3), np.int32)
            out["order"] = np.zeros(n, np.uint32)
            pts.cell = _ptr(out["cell"], C.c_int32)
            pts.order = _ptr(out["order"], C.c_uint32)
        keep = _abi.MPM_F_ALL | _abi.MPM_F_J2 | _abi.MPM_F_CELL
        check(lib().mpm_readback(self._h, mask & keep, C.byref(pts)), self._h)
        return out

    def getGridPoints(self):
        """GridPoint fields of every node (PhysicsEngine.h:65, GridPoint.h:12-28),
        x fastest: ``active``, ``fixed`` (FIX_* bits), ``x``, ``mass``, ``v``,
        ``f``, ``f_temp`` (see mpm_grid_nodes for which step they describe)."""
        nn = self.getCount_GridPoint()
        out = dict(active=np.zeros(nn, np.uint8), fixed=np.zeros(nn, np.uint8), x=np.zeros((nn, 3)),
                   mass=np.zeros(nn), v=np.zeros((nn, 3)), f=np.zeros((nn, 3)), f_temp=np.zeros((nn, 3)),
                   kernel=np.zeros(nn), kernel_grad=np.zeros((nn, 3)))
        nd = MpmNodes()
        nd.n = nn
        for k, a in out.items():
            setattr(nd, k, _ptr(a, C.c_uint8) if a.dtype == np.uint8 else _ptr(a))
        check(lib().mpm_grid_nodes(self._h, C.byref(nd)), self._h)
        return out

    def getGridPoints_Kernel(self):
        """The kernel grid of the contact step (PhysicsEngine.h:66): node
        positions with ``kernel`` (d_Kernel) and ``kernel_grad``
        (d3_Kernel_Gradient) of the last contact step (zeros before one)."""
        g = self.getGridPoints()
        return {k: g[k] for k in ("x", "kernel", "kernel_grad", "active", "fixed")}

    def checks(self):
        out = np.zeros(12)
        check(lib().mpm_checks(self._h, _ptr(out)), self._h)
        keys = ["grid_mass", "reaction_y", "kinetic_2", "px", "py", "pz", "nonfinite", "time", "cycle",
                "rebins", "bricks", "grid_nodes"]
        return dict(zip(keys, out.tolist()))

    def reaction_history(self, cap=1 << 20):
        buf = np.zeros(cap)
        n = C.c_int32(0)
        check(lib().mpm_reaction_history(self._h, _ptr(buf), cap, C.byref(n)), self._h)
        return buf[: n.value].copy()

    def group_reaction_history(self, cap=1 << 20):
        """Per-step reaction of every batched world, shape (steps, n_groups)."""
        ng = int(self.scene.n_groups or 0)
        if ng == 0:
            return np.zeros((0, 0))
        buf = np.zeros(cap * ng)
        n = C.c_int32(0)
        check(lib().mpm_group_reaction_history(self._h, _ptr(buf), cap, C.byref(n)), self._h)
        return buf[: n.value * ng].reshape(n.value, ng).copy()

    def kernel_times(self):
        out = np.zeros(6)
        check(lib().mpm_kernel_times(self._h, _ptr(out)), self._h)
        return dict(particle_ms=out[0], particle_launches=int(out[1]), grid_ms=out[2],
                    grid_launches=int(out[3]), rebin_ms=out[4], rebins=int(out[5]))

    def getTime_Current(self):
        t = C.c_double()
        check(lib().mpm_get_time(self._h, C.byref(t), None, None), self._h)
        return t.value

    def getTime_End(self):
        return self.scene.time_end

    def getTime_Increment(self):
        return self.scene.dt_max

    def reportConsole(self, sDescription=""):
        """The reference's console line for the current state (printed when
        the engine is verbose, appended to the log file of ``set_log``)."""
        n = C.c_int32()
        cap = 4096 + 2 * len(sDescription)
        buf = C.create_string_buffer(cap)
        check(lib().mpm_report(self._h, sDescription.encode(), buf, cap, C.byref(n)), self._h)
        return buf.value.decode()

    def set_log(self, path):
        """Append every report line to ``path`` (the reference's ./bmp/log.txt);
        None stops writing."""
        check(lib().mpm_set_log(self._h, None if path is None else str(path).encode()), self._h)

    def getTime_Runtime(self):
        ms = C.c_double()
        check(lib().mpm_last_step_ms(self._h, C.byref(ms)), self._h)
        return ms.value * 1e-3

    def getCount_MaterialPoint(self):
        return self.scene.n

    def getCount_GridPoint(self):
        nn = self.scene.nodes
        return nn[0] * nn[1] * nn[2]

    # ------------------------------------------------------------ text dumps
    def getScript_MaterialPoints(self):
        """MaterialPoint_Factory::getScript (MaterialPoint_Factory_getScript.cpp:4-28)
        over every point at the current state, header first."""
        return dump.material_points_script(self.getMaterialPoints(_abi.MPM_F_X)["x"], self.scene.mass)

    def getScript_GridPoints(self, active_only=True):
        """GridPoint_Factory::getScript (GridPoint_Factory.h:69-100) over the
        grid nodes (those with mass only, by default), header first; node mass
        from mpm_grid_nodes (the P2G mass of the points' current positions)."""
        return dump.grid_points_script(self.scene, active_only=active_only, mass=self.getGridPoints()["mass"])

    def saveLatex(self, strPrefix, strPostfix=".txt"):
        """The points' getScript table written to ``strPrefix + strPostfix``
        (MaterialPoint_Factory::saveLatex, declared at MaterialPoint_Factory.h:41
        and never defined by the reference); returns the path."""
        return dump.save_latex(self.getMaterialPoints(_abi.MPM_F_X)["x"], self.scene.mass, strPrefix, strPostfix)


def sort_keys(keys, device=0):
    """Device stable radix sort of u32 keys -> (sorted, permutation)."""
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    s = np.zeros_like(k)
    p = np.zeros_like(k)
    check(lib().mpm_sort_keys(_ptr(k, C.c_uint32), len(k), _ptr(s, C.c_uint32), _ptr(p, C.c_uint32), device), None)
    return s, p
