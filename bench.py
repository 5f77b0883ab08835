"""Benchmark: particle-steps/s of the explicit MPM step on one MI355X.

Workload: a 10 x 10 tiling of the re-entrant polygon cell
at h = 0.1 mm, about 1.0M deformable material points, J2 hardening
(E = 70 GPa, nu = 0.33, sigma_y = 250 MPa, H = 700 MPa), top platen ramped to
1 m/s over 100 steps (the stated 30% strain in 5000 steps blows up under the
reference update, oracle included), bottom node plane fixed.  One "step" is one MPM
time step (grid update + fused G2P/constitutive/P2G) over all points, with the
state resident in HBM.

Prints ONE JSON line.  ``value`` is the FP32-state (FP64 return map / det F)
rate; the FP64 "plain" rate of the same scene on the same GPU is reported as
``fp64_value`` and their ratio as ``fp32_over_fp64``.  ``roofline`` covers the
dominant kernel (the fused particle kernel) from HIP-event timing of every
launch in a profiled pass; ``cpu_baseline`` times the FP64 CPU oracle (a
restatement of the reference step, 1 thread as the reference default) on a
bounded sample of the same scene.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 1M MPs (FP32 vs FP64 plain, same GPU); adjoint/forward ratio"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes (read once + write once per step; sort / arena / scratch
# traffic excluded): per material point and per active grid node
MP_BYTES = {("j2", 4): 232, ("j2", 8): 464, ("elastic", 4): 176, ("elastic", 8): 352}
# the fused particle kernel owns P2G node writes (7 words) and G2P node reads (7)
NODE_BYTES_PARTICLE_KERNEL = {4: 56, 8: 112}


def build_scene(scale, speed):
    from code_mpm_auxetics_amd import scene
    return scene.config2_world(material="j2", scale=scale, steps=5000, platen="ramp", speed=speed)


def run_gpu(sc, precision, steps, warmup, profile_steps, shape=None):
    import code_mpm_auxetics_amd as m
    kw = dict(brick=shape[0], drift_check=shape[1]) if shape else {}
    eng = m.PhysicsEngine(sc, precision=precision, track_strain=False, **kw)
    if warmup:
        eng.run_steps(warmup)
    t0 = time.perf_counter()
    eng.run_steps(steps)
    t1 = time.perf_counter()
    c = eng.checks()
    dev_ms = eng.getTime_Runtime() * 1e3
    eng.close()
    res = dict(wall_s=t1 - t0, device_ms=dev_ms, checks=c)
    if profile_steps:
        eng = m.PhysicsEngine(sc, precision=precision, track_strain=False, profile=True, **kw)
        eng.run_steps(warmup + steps)      # same trajectory position as the timed window
        eng.run_steps(profile_steps)
        res["kt"] = eng.kernel_times()
        res["checks_prof"] = eng.checks()
        eng.close()
    return res


def measured_traffic():
    """HBM bytes per launch of the particle kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by a 256 MiB copy), or None."""
    f = os.path.join(ROOT, "profiles", "k_particle_fp32_config2_1M.traffic.json")
    if not os.path.exists(f):
        return None, None
    with open(f) as fh:
        d = json.load(fh)
    return int(d["traffic_bytes"]), d.get("source")


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_rate(sc, nthreads, seconds, max_steps):
    from oracle.oracle import Oracle
    orc = Oracle(sc, nthreads=nthreads, strain_form=0)
    orc.run_steps(1)                       # warm-up (page in, first grid sweep)
    steps = 0
    t0 = time.perf_counter()
    while True:
        orc.run_steps(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= max_steps:
            break
    return sc.n_deformable * steps / el, steps


def cpu_baseline(sc, seconds=15.0):
    """The FP64 oracle (a restatement of the reference step) on the host:
    1 thread, the reference default (_MAX_N_THREADS 1, PhysicsEngine.h:23),
    and all of this process's CPU share (OpenMP over points with per-thread
    node grids, the reference's privatized scheme), each on a bounded sample
    of the same scene."""
    n = sc.n_deformable
    v1, s1 = _oracle_rate(sc, 1, seconds, 50)
    threads = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)))
    vt, st = _oracle_rate(sc, threads, 0.5 * seconds, 100)
    return dict(value=v1, unit="particle-steps/s", cores=1, kind="port",
                sample=f"{s1} steps of the {n}-point config-2 scene, FP64 oracle "
                       f"(restatement of the reference step), 1 thread as the reference default",
                all_cores=dict(value=vt, cores=threads, sample=f"{st} steps of the same scene, OpenMP, "
                                                               f"privatized node grids"),
                cpu_model=_cpu_model(), nproc=os.cpu_count())


def field_errors(m, sc, steps=100):
    """Relative L2 error of every reported point field after ``steps`` steps,
    for the FP32-state and the FP64 GPU runs, against the FP64 oracle (the
    CPU restatement of the reference step, all host threads) from the same
    initial state: displacement, velocity, F - I, stress, plastic strain,
    alpha.  100 steps is where the platen ramp ends; earlier, strains are
    ~1e-8 and below the resolution of FP32 state."""
    from oracle.oracle import Oracle
    threads = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)))
    orc = Oracle(sc, nthreads=threads, strain_form=1)
    for _ in range(steps):
        orc.step(sc.dt_max)
        orc.w.time += sc.dt_max
    o = orc.points()
    x0 = sc.x

    def rel(a, b):
        a = np.asarray(a, np.float64).ravel()
        b = np.asarray(b, np.float64).ravel()
        d = np.linalg.norm(b)
        return float(np.linalg.norm(a - b) / d) if d > 0 else float(np.linalg.norm(a))

    out = {"steps": steps, "vs": "FP64 oracle (CPU restatement of the reference step, L dt strain form)",
           "oracle_threads": threads}
    keep = sc.dc == 0
    for name, prec in (("fp32", m.FP32), ("fp64", m.FP64)):
        eng = m.PhysicsEngine(sc, precision=prec, track_strain=False, strain_form=1)
        eng.run_steps(steps)
        g = eng.getMaterialPoints()
        eng.close()
        out[name] = {"disp": rel((g["x"] - x0)[keep], (o["x"] - x0)[keep]), "v": rel(g["v"][keep], o["v"][keep]),
                     "F": rel(g["F"][keep] - np.array([1, 0, 0, 0, 1, 0, 0, 0, 1.0]),
                              o["F"][keep] - np.array([1, 0, 0, 0, 1, 0, 0, 0, 1.0])),
                     "stress": rel(g["stress"][keep], o["stress"][keep]),
                     "eps_p": rel(g["eps_p"][keep], o["eps_p"][keep]), "alpha": rel(g["alpha"][keep], o["alpha"][keep])}
    return out


def plastic_window(m, steps, pre=1500, sy_scale=0.06, speed=1.0):
    """The config-2 scene with its yield stress scaled by ``sy_scale`` (250 MPa
    -> 15 MPa; everything else unchanged), so that the J2 return map takes its
    FP64 plastic branch inside the stable window of the loading (the 250 MPa
    scene stays elastic until the reference update diverges, near step
    4000).  ``pre`` untimed steps reach the yielding state, then ``steps``
    steps are timed in FP32 and FP64.  Reports the fraction of points whose
    plastic strain changed during the window and the FP32 fields against the
    FP64 run from the same start."""
    sc = build_scene(0.21, speed)
    d = sc.dc == 0
    sc.sy = sc.sy.copy()
    sc.sy[d] *= sy_scale
    out = {"workload": "config2_polygon_mesh_10x10_j2_plastic_window", "sy_pa": float(sc.sy[d][0]),
           "pre_steps": pre, "steps": steps, "n_points": int(d.sum())}
    state = {}
    F = m._abi
    for name, prec in (("fp32", m.FP32), ("fp64", m.FP64)):
        with m.PhysicsEngine(sc, precision=prec, track_strain=False) as eng:
            eng.run_steps(pre)
            ep0 = eng.getMaterialPoints(F.MPM_F_EPSP)["eps_p"][d]
            t0 = time.perf_counter()
            eng.run_steps(steps)
            t1 = time.perf_counter()
            g = eng.getMaterialPoints(F.MPM_F_X | F.MPM_F_STRESS | F.MPM_F_EPSP | F.MPM_F_ALPHA)
        changed = np.any(g["eps_p"][d] != ep0, axis=1)
        out[name] = {"ms_per_step": round((t1 - t0) * 1e3 / steps, 5),
                     "value": round(int(d.sum()) * steps / (t1 - t0), 1),
                     "yield_fraction": round(float(changed.mean()), 4)}
        state[name] = g

    def rel(a, b):
        a = np.asarray(a, np.float64).ravel()
        b = np.asarray(b, np.float64).ravel()
        n = np.linalg.norm(b)
        return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a))
    a, b = state["fp32"], state["fp64"]
    out["fp32_vs_fp64"] = {"disp": rel((a["x"] - sc.x)[d], (b["x"] - sc.x)[d]), "stress": rel(a["stress"][d], b["stress"][d]),
                           "eps_p": rel(a["eps_p"][d], b["eps_p"][d]), "alpha": rel(a["alpha"][d], b["alpha"][d])}
    return out


def config1_run(m, out_dir=None):
    """Config 1 as stated: one 10 x 10 x 1 mm polygon cell, 0.8 mm walls,
    h = 0.125 mm, elastic, platen at 0.1 m/s, 2000 FP64 steps.  Writes the
    per-step platen force and the final positions and stresses (the config's
    outputs) and compares them with the oracle's run frozen in
    tests/golden/config1_output.npz."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_golden
    sc = make_golden.config1_scene()
    z = np.load(os.path.join(ROOT, "tests", "golden", "config1_output.npz"))
    steps = int(z["steps"])
    with m.PhysicsEngine(sc, precision=m.FP64, track_strain=False) as eng:
        t0 = time.perf_counter()
        eng.runSimulation_Classic_SinglePass_MP(steps * sc.dt_max)
        t1 = time.perf_counter()
        force = eng.reaction_history()
        g = eng.getMaterialPoints(m._abi.MPM_F_X | m._abi.MPM_F_STRESS | m._abi.MPM_F_V)
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        np.savetxt(os.path.join(out_dir, "config1_platen_force.csv"),
                   np.column_stack([np.arange(1, len(force) + 1) * sc.dt_max, force]), delimiter=",",
                   header="time_s,force_y_N", comments="")
        np.savez_compressed(os.path.join(out_dir, "config1_final_points.npz"), x=g["x"], stress=g["stress"],
                            dc=sc.dc)
    idx = z["idx"]

    def rel(a, b):
        return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))
    return {"workload": "config1_polygon_cell_elastic_fp64", "n_points": sc.n_deformable, "steps": len(force),
            "ms_per_step": round((t1 - t0) * 1e3 / len(force), 5), "dtype": "f64",
            "value": round(sc.n_deformable * len(force) / (t1 - t0), 1), "unit": "particle-steps/s",
            "final_platen_force_N": float(force[-1]),
            "vs_oracle_fixture": {"force": rel(force, z["reaction"]), "disp": rel(g["x"][idx] - sc.x[idx], z["x"] - sc.x[idx]),
                                  "stress": rel(g["stress"][idx], z["stress"])},
            "outputs": out_dir}


def batched_run(m, steps, warmup):
    """Config 4: 256 independent cells (SplitMix64 seed 20240607) packed into
    one grid, FP32, one launch per step for all cells (and, timed beside it,
    the persistent launch: a resident grid of blocks pulling every cell's
    bricks from per-XCD work queues); aggregate particle-steps/s and the
    per-cell effective Poisson's ratio."""
    from code_mpm_auxetics_amd import scene
    ws = scene.config4_worlds()
    sc = scene.batch_worlds(ws)
    n = sc.n_deformable
    times = {}
    for persistent in (True, False):
        with m.PhysicsEngine(sc, precision=m.FP32, track_strain=False, persistent=persistent) as eng:
            eng.run_steps(warmup)
            t0 = time.perf_counter()
            eng.run_steps(steps)
            t1 = time.perf_counter()
            x1 = eng.getMaterialPoints(m._abi.MPM_F_X)["x"]
        times[persistent] = t1 - t0
    nus = []
    for i in range(len(ws)):
        keep = (sc.info["world"] == i) & (sc.dc == 0)
        nus.append(scene.effective_poisson(sc.x, x1, keep))
    nus = np.array(nus)
    poly = np.array([w.info["topology"] == "polygon" for w in ws])
    t = times[False]
    print(json.dumps({"workload": "config4_batched_256_cells", "n_points": n, "cells": len(ws),
                      "grid_cells": list(sc.cells), "steps": steps, "warmup": warmup,
                      "value": round(n * steps / t, 1), "unit": "particle-steps/s",
                      "ms_per_step": round(t * 1e3 / steps, 4), "dtype": "f32",
                      "persistent_value": round(n * steps / times[True], 1),
                      "persistent_ms_per_step": round(times[True] * 1e3 / steps, 4),
                      "poisson_median_polygon": round(float(np.median(nus[poly])), 4),
                      "poisson_median_swisscheese": round(float(np.median(nus[~poly])), 4),
                      "poisson_min": round(float(nus.min()), 4), "poisson_max": round(float(nus.max()), 4),
                      "simulated_time_s": steps * sc.dt_max + warmup * sc.dt_max}), flush=True)


def config3_run(m, steps, warmup):
    """Config 3: the swiss-cheese block (8 x 8 holes, Neo-Hookean, about 2.4M
    points) under a point platen on the contact path (two-field grid, kernel
    grid, surface normals, contact correction with Coulomb friction 0.2), FP32
    state with FP64 checks.  Times ``steps`` steps after ``warmup``."""
    from code_mpm_auxetics_amd import scene
    sc = scene.config3_world()
    n = sc.n_deformable
    eng = m.PhysicsEngine(sc, precision=m.FP32, track_strain=False, friction=0.2)
    eng.run_steps(warmup, path=m.PATH_CONTACT)
    t0 = time.perf_counter()
    eng.run_steps(steps, path=m.PATH_CONTACT)
    t1 = time.perf_counter()
    c = eng.checks()
    eng.close()
    return {"workload": "config3_swisscheese_block_contact", "n_points": n, "grid_cells": list(sc.cells),
            "steps": steps, "warmup": warmup, "value": round(n * steps / (t1 - t0), 1), "unit": "particle-steps/s",
            "ms_per_step": round((t1 - t0) * 1e3 / steps, 4), "dtype": "f32", "friction": 0.2,
            "dt_s": sc.dt_max, "kinetic_2": c["kinetic_2"], "reaction_y": c["reaction_y"],
            "nonfinite": c["nonfinite"]}


def config5_run(reps=3):
    """Config 5: the differentiable run on the 4x4 polygon mesh (about 160k
    points, per-point E log-uniform in [0.5, 2] GPa), 400 steps with
    checkpoints every 20, FP32 state as the config states (FP64 arithmetic
    inside each step and in the adjoint) and, beside it, FP64 state.  Times
    the forward alone and forward + adjoint through the
    torch.autograd.Function with the effective-Poisson loss (best of
    ``reps``, wall clock around each call, state on the GPU)."""
    import torch
    import code_mpm_auxetics_amd as m
    from code_mpm_auxetics_amd import diff, scene
    sc = scene.config5_world()
    sets = diff.boundary_sets(sc)
    res = {}
    for name, prec in (("fp32", m.FP32), ("fp64", m.FP64)):
        with diff.DiffRun(sc, 400, checkpoint=20, precision=prec) as run:
            run.forward(sc.E, sc.x)
            t_f, t_fb = [], []
            for _ in range(reps):
                t0 = time.perf_counter()
                u = run.forward(sc.E, sc.x)
                t_f.append(time.perf_counter() - t0)
            for _ in range(reps):
                t0 = time.perf_counter()
                E = torch.tensor(sc.E, requires_grad=True)
                x0 = torch.tensor(sc.x, requires_grad=True)
                loss = diff.poisson_loss(x0, diff.rollout(run, E, x0), sets)
                loss.backward()
                t_fb.append(time.perf_counter() - t0)
        res[name] = dict(f=min(t_f), fb=min(t_fb), loss=float(loss.detach()),
                         nu=float(diff.effective_poisson(sc.x, u, sets)))
    # the same 400 steps on the product step (no checkpoints or node slots
    # kept), FP64 and FP32 state, L dt strain increment as in the run above
    prod = {}
    for name, prec in (("fp64", m.FP64), ("fp32", m.FP32)):
        best = None
        for _ in range(reps):
            with m.PhysicsEngine(sc, precision=prec, track_strain=False, strain_form=1) as eng:
                t0 = time.perf_counter()
                eng.run_steps(400)
                el = time.perf_counter() - t0
            best = el if best is None else min(best, el)
        prod[name] = best
    r32, r64 = res["fp32"], res["fp64"]
    return {"workload": "config5_polygon_mesh_4x4_elastic_adjoint", "n_points": sc.n_deformable, "steps": 400,
            "checkpoint": 20, "dtype": "f32", "state": "fp32 (u, v, F - I, sigma, alpha); fp64 step and adjoint arithmetic",
            "forward_ms": round(r32["f"] * 1e3, 2), "forward_adjoint_ms": round(r32["fb"] * 1e3, 2),
            "adjoint_over_forward": round(r32["fb"] / r32["f"], 3),
            "forward_particle_steps_per_s": round(sc.n_deformable * 400 / r32["f"], 1),
            "loss": r32["loss"], "nu_eff": r32["nu"],
            "fp64_state": {"forward_ms": round(r64["f"] * 1e3, 2), "forward_adjoint_ms": round(r64["fb"] * 1e3, 2),
                           "adjoint_over_forward": round(r64["fb"] / r64["f"], 3), "loss": r64["loss"]},
            "product_forward_fp64_ms": round(prod["fp64"] * 1e3, 2),
            "product_forward_fp32_ms": round(prod["fp32"] * 1e3, 2),
            "forward_adjoint_over_product_fp32_forward": round(r32["fb"] / prod["fp32"], 3),
            "forward_adjoint_over_product_fp64_forward": round(r64["fb"] / prod["fp64"], 3)}


def counters_run(m, sc, warmup, steps):
    """FP32 steps of the bench scene, then one plain device-to-device copy of
    256 MiB (the calibration kernel for FETCH_SIZE / WRITE_SIZE)."""
    import torch
    eng = m.PhysicsEngine(sc, precision=m.FP32, track_strain=False)
    eng.run_steps(warmup + steps)
    c = eng.checks()
    eng.close()
    a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"counters_run": True, "steps": warmup + steps, "n_points": sc.n_deformable,
                      "grid_nodes": int(c["grid_nodes"]), "copy_bytes": 2 * a.numel() * 4}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--scale", type=float, default=0.21)
    ap.add_argument("--speed", type=float, default=1.0, help="platen speed, m/s (ramped over 100 steps)")
    ap.add_argument("--fp64-steps", type=int, default=200)
    ap.add_argument("--profile-steps", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="config2", choices=["config1", "config2", "config3", "config4", "config5"],
                    help="config4: the 256-cell batch (~8M points) in one launch per step; config5: the "
                         "differentiable run alone; each prints its own line")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the config-5 adjoint timing in the main line")
    ap.add_argument("--no-plastic", action="store_true", help="skip the plastic-window leg in the main line")
    ap.add_argument("--out-dir", default=None, help="config1: directory for the force history and final points")
    ap.add_argument("--tune", action="store_true",
                    help="pick brick size and batch length with the autotuner before timing")
    ap.add_argument("--counters", action="store_true",
                    help="short FP32 run plus a 256 MiB device copy, for rocprofv3 --pmc passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is a design "
                                   "(see DESIGN.md), not a measured path", "n_gpus": args.gpus}))
        sys.exit(2)
    import code_mpm_auxetics_amd as m

    if args.workload == "config1":
        print(json.dumps(config1_run(m, args.out_dir)), flush=True)
        return
    if args.workload == "config4":
        batched_run(m, args.steps, args.warmup)
        return
    if args.workload == "config5":
        print(json.dumps(config5_run()), flush=True)
        return
    if args.workload == "config3":
        print(json.dumps(config3_run(m, args.steps, args.warmup)), flush=True)
        return
    sc = build_scene(args.scale, args.speed)
    n = sc.n_deformable
    if args.counters:
        counters_run(m, sc, args.warmup, args.steps)
        return
    shape32 = shape64 = None
    if args.tune:
        from code_mpm_auxetics_amd import autotune
        t32 = autotune.tune(sc, precision=m.FP32)
        t64 = autotune.tune(sc, precision=m.FP64)
        shape32, shape64 = (t32["brick"], t32["drift_check"]), (t64["brick"], t64["drift_check"])
    r32 = run_gpu(sc, m.FP32, args.steps, args.warmup, args.profile_steps, shape32)
    r64 = run_gpu(sc, m.FP64, args.fp64_steps, args.warmup, 0, shape64)
    v32 = n * args.steps / r32["wall_s"]
    v64 = n * args.fp64_steps / r64["wall_s"]

    kt = r32["kt"]
    avg_ms = kt["particle_ms"] / max(kt["particle_launches"], 1)
    avg_grid_ms = kt["grid_ms"] / max(kt["grid_launches"], 1)
    nodes = r32["checks_prof"]["grid_nodes"]
    bytes_launch = n * MP_BYTES[("j2", 4)] + nodes * NODE_BYTES_PARTICLE_KERNEL[4]
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic()
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
            "kernel": "k_particle<float,true>", "avg_launch_ms": round(avg_ms, 5),
            "algorithmic_bytes_per_launch": int(bytes_launch)}
    step_bytes = n * MP_BYTES[("j2", 4)] + nodes * 108
    out = {
        "metric": METRIC, "value": round(v32, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(r32["wall_s"] * 1e3 / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "config2_polygon_mesh_10x10_j2", "n_points": n, "h_mm": 0.1,
                   "grid_cells": list(sc.cells), "precision": "fp32 state, fp64 return map and det F",
                   "platen_speed_m_s": round(sc.info["platen_speed"], 3), "platen_ramp_steps": 100,
                   "dt_s": sc.dt_max,
                   "parallelism": "single GPU",
                   "launch": {"fp32": list(shape32) if shape32 else "default (bricks 4x4x10, batch 8)",
                              "fp64": list(shape64) if shape64 else "default (bricks 4x4x10, batch 8)",
                              "tuned": bool(args.tune)}},
        "fp64_value": round(v64, 1), "fp64_ms_per_step": round(r64["wall_s"] * 1e3 / args.fp64_steps, 5),
        "fp32_over_fp64": round(v32 / v64, 3),
        "step_roofline_frac": round(step_bytes * args.steps / r32["wall_s"] / 1e9 / HBM_PEAK_GBS, 4),
        "grid_kernel_avg_ms": round(avg_grid_ms, 5),
        "rebin_ms_total_profiled": round(kt["rebin_ms"], 3), "rebins_profiled": kt["rebins"],
        "active_grid_nodes": int(nodes),
        "roofline": roof,
    }
    if not args.no_plastic:
        out["plastic_window"] = plastic_window(m, args.steps)
    if not args.no_adjoint:
        c5 = config5_run()
        out["adjoint_over_forward"] = c5["adjoint_over_forward"]
        out["adjoint"] = c5
    if not args.no_cpu:
        out["field_errors"] = field_errors(m, sc)
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
