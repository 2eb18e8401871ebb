#!/usr/bin/env python3
"""Benchmark: sphere BVE direct-sum RK4 steps on one MI355X.

Workload (N = 655k): icosahedral mesh at level 8, 655,362 vertex particles plus
1,310,720 panel-centre particles (T = 1,966,082 targets, S = 1,310,720 active
sources), relative vorticity = 4 pi exp(-16 |x - x0|^2) (x0 at lat 0.3, lon 0)
plus a Rossby-Haurwitz-4 background, Omega = 2 pi, dt = 0.01.  One step is a
classical RK4 step on the device (4 Biot-Savart sums of T*S - S pairs, FP32
pair arithmetic with compensated accumulation, FP64 stage algebra, stage
points projected onto the sphere in FP64).

Prints one JSON line.  `value` = G pair-interactions/s over the timed steps.
Extra fields: RK4 steps/s, velocity error of the FP32 sum against the FP64 GPU
sum over every target, the pair kernel's roofline, the CPU baseline (lpm-v2's
own Fortran loop, built from its sources, on one MPI rank per host core over a
bounded block of targets; beside it the FP64 oracle on the same cores) and
`learned_far_field`: the learned far-field form on the same workload (level-3
cells, trained on 1e7 direct-sum samples of its own run, 100 RK4 steps judged
by the plain sum every 10 steps), its speedup over the plain steps and the
largest judged error.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lpm-v2_amd"))
sys.path.insert(0, ROOT)

METRIC = "G pair-interactions/s & RK4 steps/s at N=655k; speedup vs plain at stated err"
# sphere Biot-Savart flop contract (+, -, x = 1, FMA = 2, rcp = 1, per-source and
# per-target invariants hoisted): x.y 5, R^2 - x.y 1, 1/den 1, x w_j 1, cross
# product 9, accumulate 3 x 2 = 23 flops per pair
FLOPS_PER_PAIR = 23
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# HBM bytes per pair-kernel launch at level 8 from rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs, scaled by the same counters on a 1 GiB device copy (factors
# 2.0 and 1.0): profiles/sphere_bs_fp32_icos8_hbm_counters.md
MEASURED_TRAFFIC_BYTES = {8: 4.22e9}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(level: int):
    from lpmgpu import ic, mesh

    m = mesh.icosahedral_mesh(level, cell_level=3)  # level-3 faces: the far-field cells
    x0 = ic.lat_lon_point(0.3, 0.0)
    zeta = ic.gaussian_vortex(m.x, m.y, m.z, x0, 4.0 * np.pi, 16.0) + ic.rh4_vorticity(m.x, m.y, m.z)
    return m, zeta


def cpu_baseline(m, zeta, target_seconds: float, max_threads: int):
    """FP64 oracle (the reference loop restated in C, OpenMP over a contiguous
    target block) on a bounded block of targets; returns G pairs/s."""
    from oracle import oracle

    threads = oracle.set_threads(max_threads)
    ns = int(m.active.sum())
    b = 256
    rate = None
    t_total = 0.0
    while True:
        t0 = time.perf_counter()
        oracle.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, 1.0, 0, b)
        dt = time.perf_counter() - t0
        t_total += dt
        pairs = b * ns - int(m.active[:b].sum())
        rate = pairs / dt
        if dt > 0.25 * target_seconds or t_total > target_seconds or b >= m.n:
            break
        b = min(m.n, int(b * max(2.0, 0.5 * target_seconds / max(dt, 1e-3))))
    out = {
        "value": rate / 1e9,
        "unit": "G pair-interactions/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP64 oracle BVESphereVelocity loop, targets [0,{b}) x {ns} active sources "
                  f"({pairs / 1e9:.2f} G pairs, {dt:.1f} s)",
        "cpu_model": host_cpu_model(),
        "host": host_description(),
    }
    # the same loop on one core, over a block sized to about a tenth of the budget
    oracle.set_threads(1)
    b1 = max(16, min(m.n, int(0.1 * target_seconds * rate / threads / max(ns, 1))))
    t0 = time.perf_counter()
    oracle.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, 1.0, 0, b1)
    dt1 = time.perf_counter() - t0
    pairs1 = b1 * ns - int(m.active[:b1].sum())
    out["one_core"] = {"value": pairs1 / dt1 / 1e9, "sample": f"targets [0,{b1}) ({pairs1 / 1e9:.2f} G pairs, "
                                                              f"{dt1:.1f} s)"}
    oracle.set_threads(max_threads)
    return out


REF_CPU_BENCH = os.path.join(ROOT, "oracle", "_ref", "ref_cpu_bench")
MPIEXEC = "/opt/conda/bin/mpiexec"


def reference_cpu_baseline(m, zeta, target_seconds: float, ranks: int):
    """lpm-v2's own Fortran loop (SetVelocityOnMesh, SphereBVE.f90:489-531),
    compiled from its sources by oracle/ref.mk, on `ranks` MPI ranks (one per
    core) over a bounded block of targets sized to about target_seconds, as
    the reference runs it.  Its block is checked against the FP64 oracle.
    Returns None where the driver or MPI launcher is not there."""
    import shutil
    import subprocess
    import tempfile

    if not (os.path.exists(REF_CPU_BENCH) and os.path.exists(MPIEXEC)):
        return None
    from oracle import oracle

    tmp = tempfile.mkdtemp(prefix="lpm_ref_cpu_")
    try:
        for k, a in (("x", m.x), ("y", m.y), ("z", m.z), ("area", m.area), ("active", m.active), ("relvort", zeta)):
            np.asarray(a, dtype="<f8").tofile(os.path.join(tmp, f"in_{k}.f64"))
        cmd = [MPIEXEC, "-launcher", "fork", "-n", str(ranks), REF_CPU_BENCH, tmp, str(m.n), str(target_seconds)]
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=4 * target_seconds + 120)
        except subprocess.TimeoutExpired:
            log("[cpu reference] timed out")
            return None
        if r.returncode != 0:
            log(f"[cpu reference] exit {r.returncode}: {(r.stdout + r.stderr)[-400:]}")
            return None
        vals = None
        for line in reversed(r.stdout.splitlines()):  # the driver's result line: 5 numbers
            try:
                nr, k, ns, sec, rate = line.split()
                vals = int(nr), int(k), int(ns), float(sec), float(rate)
                break
            except ValueError:
                continue
        if vals is None:
            log(f"[cpu reference] no result line in: {r.stdout[-400:]}")
            return None
        nr, k, ns, sec, rate = vals
        got = np.stack([np.fromfile(os.path.join(tmp, f"out_{c}.f64"), dtype="<f8") for c in "uvw"])
        ref = np.stack([np.asarray(c)[:k] for c in oracle.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active,
                                                                               1.0, 0, k)])
        diff = float(np.abs(got - ref).max() / np.abs(ref).max())
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return {
        "value": rate / 1e9,
        "unit": "G pair-interactions/s",
        "cores": int(nr),
        "kind": "reference",
        "sample": f"lpm-v2's SetVelocityOnMesh (amdflang -O2, MPICH), {nr} ranks, targets [0,{k}) x {ns} active "
                  f"sources ({k * ns / 1e9:.2f} G pairs, {sec:.1f} s); its block agrees with the FP64 oracle "
                  f"to {diff:.1e}",
        "max_rel_diff_vs_oracle": diff,
    }


def velocity_error_fp64(ctx, st: dict, m) -> dict:
    """Normwise max, RMS and pointwise errors of the stored FP32 velocity of a
    solver state against the FP64 GPU direct sum of the same state."""
    g64 = np.array(ctx.bve_sphere_velocity(st["x"], st["y"], st["z"], st["relvort"], m.area, m.active,
                                           precision="fp64"))
    g32 = np.array([st["u"], st["v"], st["w"]])
    d = np.sqrt(((g32 - g64) ** 2).sum(0))
    mag = np.sqrt((g64**2).sum(0))
    return {
        "max_rel": float(d.max() / mag.max()),
        "rms_rel": float(np.sqrt((d**2).sum() / (mag**2).sum())),
        "pointwise_max_rel": float((d / np.maximum(mag, 1e-3 * mag.max())).max()),
        "targets": int(d.shape[0]),
    }


def host_description() -> dict:
    """Host CPUs as the process sees them: affinity set, nproc, model string,
    family/model numbers (which identify the core generation where a
    hypervisor shortens the name) and lscpu's model name."""
    import subprocess

    info = {"affinity_cpus": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                k, _, v = line.partition(":")
                k = k.strip()
                if k in ("model name", "cpu family", "model", "stepping", "cpu MHz") and k not in info:
                    info[k.replace(" ", "_")] = v.strip()
                if k == "flags":
                    break
    except OSError:
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info["lscpu_" + k.strip().lower().replace("(s)", "s").replace(" ", "_")] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def default_cpu_threads() -> int:
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(cap))) if cap and cap.isdigit() else n


def host_cpu_model() -> str:
    """The host CPU's model name from /proc/cpuinfo (empty if unreadable)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def cfg_cpu_baseline(workload: str, res: dict, target_seconds: float, max_threads: int):
    """The FP64 oracle (the reference loops restated in C, OpenMP over contiguous
    target blocks) beside a cfg workload, on the box's host cores.  cfg0 runs
    whole (its 10 RK4 steps); cfg1 and cfg3 time the sphere sum on a bounded
    target block and extrapolate a step (4 sums); cfg2 times the PSE sums on a
    bounded block and extrapolates one evaluation.  Adds res["cpu_baseline"]."""
    import types

    from lpmgpu import workloads
    from oracle import oracle

    threads = oracle.set_threads(max_threads)
    if workload == "cfg0":
        w = workloads.cfg0()
        m = w["mesh"]
        u, v, z = oracle.bve_sphere_velocity(m.x, m.y, m.z, w["relvort"], m.area, m.active)
        st = dict(x=m.x, y=m.y, z=m.z, relvort=w["relvort"], u=u, v=v, w=z)
        t0 = time.perf_counter()
        oracle.bve_timestep(st, m.area, m.active, 1.0, w["omega"], w["dt"], nsteps=w["steps"])
        ms = 1e3 * (time.perf_counter() - t0) / w["steps"]
        out = {"ms_per_step": round(ms, 3), "unit": "ms per RK4 step", "cores": threads, "kind": "port",
               "sample": f"the whole workload: {w['steps']} FP64 RK4 steps (4 sums each) at icosahedral L4"}
        if res.get("ms_per_step"):
            out["gpu_speedup"] = round(ms / res["ms_per_step"], 1)
    elif workload in ("cfg1", "cfg3"):
        if workload == "cfg1":
            w = workloads.cfg1()
            m, zeta = w["mesh"], w["relvort"]
            pairs_step = 4 * (m.n * m.n_active - m.n_active)
        else:
            probs = workloads.cfg3()["problems"]
            p = max(probs, key=lambda q: q["x"].shape[0])
            m = types.SimpleNamespace(x=p["x"], y=p["y"], z=p["z"], area=p["area"], active=p["active"],
                                      n=p["x"].shape[0])
            zeta = p["relvort"]
            pairs_step = 4 * res["pairs_per_eval"]
        out = cpu_baseline(m, zeta, target_seconds, max_threads)
        ms = pairs_step / (out["value"] * 1e9) * 1e3
        out["ms_per_step_extrapolated"] = round(ms, 1)
        if res.get("ms_per_step"):
            out["gpu_speedup"] = round(ms / res["ms_per_step"], 1)
    elif workload == "cfg2":
        w = workloads.cfg2(512)
        x, y, area, eps, f = w["x"], w["y"], w["area"], w["eps"], w["f"]
        n = x.shape[0]
        act = np.ones(n, np.int32)
        out = {"unit": "G pair-interactions/s", "cores": threads, "kind": "port"}
        for order in (2, 8):
            b = 256
            while True:
                t0 = time.perf_counter()
                oracle.pse_plane_laplacian(x, y, f, area, act, eps, order, 0, b)
                dt = time.perf_counter() - t0
                if dt > 0.2 * target_seconds or b >= n:
                    break
                b = min(n, int(b * max(2.0, 0.4 * target_seconds / max(dt, 1e-3))))
            rate = b * n / dt
            entry = {"value": rate / 1e9, "sample": f"targets [0,{b}) x {n} sources ({b * n / 1e9:.2f} G pairs, "
                                                    f"{dt:.1f} s)",
                     "ms_per_evaluation_extrapolated": round(n * n / rate * 1e3, 1)}
            k = (res.get(f"pse{order}") or {}).get("pair_kernel") or {}
            if k.get("avg_ms"):
                entry["gpu_speedup"] = round(entry["ms_per_evaluation_extrapolated"] / k["avg_ms"], 1)
            out[f"pse{order}"] = entry
    else:
        return
    oracle.set_threads(max_threads)
    res["cpu_baseline"] = out
    log(f"[cpu {workload}] {out}")


def pse_ops_cpu_baseline(res: dict, max_threads: int, plane_block: int = 2048, sphere_block: int = 1024):
    """FP64 oracle (the reference loops restated in C, OpenMP over a contiguous
    target block) for each operator of the pse_ops workload, on the same inputs
    and a bounded block of targets; adds a cpu_baseline entry (G pairs/s) to
    each operator's result."""
    from lpmgpu import mesh, workloads
    from oracle import oracle

    threads = oracle.set_threads(max_threads)
    w = workloads.cfg2(512)
    x, y, area, eps, f = w["x"], w["y"], w["area"], w["eps"], w["f"]
    n = x.shape[0]
    act = np.ones(n, np.int32)
    u, v = np.sin(2 * y) + x, np.cos(x) * y
    r2 = x * x + y * y
    vort, div = 3.0 * np.exp(-4 * r2), 0.5 * np.sin(2 * x) * np.cos(y)
    hgt, bottom = 1.0 + 0.1 * np.exp(-6 * ((x - 0.2) ** 2 + y * y)), 0.05 * x - 0.02 * y * y
    rng = np.random.default_rng(20261202)
    px, py = rng.uniform(-0.7, 0.7, n), rng.uniform(-0.7, 0.7, n)
    m = mesh.icosahedral_mesh(7)
    seps = float(m.max_edge_length ** 0.75)
    S = int(m.active.sum())
    fs = np.sin(3 * m.z) + m.x
    g = (m.y, -m.x, 0.0 * m.z)
    lats = -0.5 * np.pi + np.arange(181) * np.pi / 180.0
    lons = np.arange(360) * np.pi / 180.0
    LA, LO = np.meshgrid(lats, lons, indexing="ij")
    qx, qy, qz = (np.cos(LO) * np.cos(LA)).ravel(), (np.sin(LO) * np.cos(LA)).ravel(), np.sin(LA).ravel()
    bp, bs = plane_block, sphere_block
    cases = {
        "plane_gradient": (lambda: oracle.pse_plane_gradient(x, y, f, area, act, eps, 0, bp), bp * n),
        "plane_double_dot": (lambda: oracle.pse_plane_double_dot(x, y, u, v, area, act, eps, 0, bp), bp * n),
        "plane_interpolation": (lambda: oracle.pse_plane_interpolate(x, y, f, area, act, eps, px[:bp], py[:bp]),
                                bp * n),
        "swe_plane_rhs": (lambda: oracle.swe_plane_rhs(x, y, vort, div, hgt, bottom, area, act, eps, 0, bp),
                          bp * (n - 1)),
        "sphere_gradient": (lambda: oracle.pse_sphere_gradient(m.x, m.y, m.z, fs, m.area, m.active, seps, 1.0, 0, bs),
                            bs * S),
        "sphere_divergence": (lambda: oracle.pse_sphere_divergence(m.x, m.y, m.z, *g, m.area, m.active, seps, 1.0,
                                                                   0, bs), bs * S),
        "sphere_interpolation_latlon": (lambda: oracle.pse_sphere_interpolate(
            m.x, m.y, m.z, fs, m.area, m.active, seps, qx[:bs], qy[:bs], qz[:bs]), bs * S),
    }
    for name, (call, pairs) in cases.items():
        t0 = time.perf_counter()
        call()
        dt = time.perf_counter() - t0
        entry = {"value": pairs / dt / 1e9, "unit": "G pair-interactions/s", "cores": threads, "kind": "port",
                 "sample": f"FP64 oracle, first {bp if name.startswith(('plane', 'swe')) else bs} targets "
                           f"({pairs / 1e9:.2f} G pairs, {dt:.1f} s)"}
        if name in res:
            res[name]["cpu_baseline"] = entry
            k = res[name].get("pair_kernel") or {}
            if k.get("avg_ms"):
                res[name]["gpu_vs_cpu"] = round(res[name]["pairs"] / (k["avg_ms"] * 1e-3) / 1e9 / entry["value"], 1)
        log(f"[cpu {name}] {entry}")


def learned_far_field(ctx, m, zeta, omega, dt, plain_ms_per_step, steps, samples, epochs, plain_end=None,
                      plain_end_step=0):
    """The learned far-field form on the same workload, as configured: 1,280
    level-3 cells, the near field within 0.1 rad summed directly (FP32
    compensated), every far cell by its 4th-order expansion in its moments
    (beyond 0.6 rad four sibling far cells as one expansion of their union), and
    the network correcting the expansion's remainder for the cells within
    0.2 rad.  Trained on direct-sum samples of the method's own run, then
    `steps` RK4 steps with the plain sum judging every 10 steps (fallback above
    1e-4) and fine-tuning the network on each judged state.  When the plain
    FP32 run's state after `plain_end_step` steps is given, the learned run's
    state after as many steps is compared with it (trajectory error)."""
    from lpmgpu import _lib as L
    from lpmgpu.farfield import EXPANSION, LEARNED, PLAIN, FarFieldBVE

    def rel(a, b):
        a, b = np.array(a), np.array(b)
        return float(np.sqrt(((a - b) ** 2).sum(0)).max() / np.sqrt((b**2).sum(0)).max())

    # band 0.2 with 20 epochs on samples drawn as the nearest of 4 band cells:
    # against band 0.3 (nearest of 7, 10 epochs) 19.3 instead of 45.6 band
    # cells per target cell, 144 instead of 163 ms per step, the largest judged
    # error 4.9e-5 instead of 5.4e-5
    # four sibling far cells as one expansion beyond 0.6 rad: 462 instead of
    # 1,258 expansions per target, 124.5 instead of 136.9 ms per step, the
    # largest judged error 4.5e-5 instead of 5.4e-5 in the same run
    near, band, order, draws, parent = 0.10, 0.20, 4, 4, 0.6
    f = FarFieldBVE(ctx, m, zeta, rotation_rate=omega, near_radius=near, band_radius=band, order=order,
                    sample_draws=draws, parent_radius=parent)
    ctx.reset_kernel_stats()
    p = f.velocity(PLAIN)
    _, ms_plain = ctx.kernel_stats(L.K_SPHERE_BS)
    ctx.reset_kernel_stats()
    e_exp = rel(f.velocity(EXPANSION), p)
    _, ms_exp = ctx.kernel_stats(6)
    t0 = time.perf_counter()
    loss = f.pretrain(dt, states=5, every=5, samples=samples, epochs=epochs)
    train_s = time.perf_counter() - t0
    ctx.reset_kernel_stats()
    e_net = rel(f.velocity(LEARNED), p)
    _, ms_net = ctx.kernel_stats(6)
    nc, bc = f.cells()
    log(f"[bench] far field: trained in {train_s:.1f}s (loss {loss:.3e}); t=0 error {e_exp:.3e} without the "
        f"network, {e_net:.3e} with it; one evaluation {ms_plain:.1f} ms plain, {ms_net:.1f} ms learned")
    # fine-tune on the judged state only where the judge finds the error above
    # half the tolerance: at band 0.2 the largest judged error was the same with
    # fine-tuning at every judge and without any (4.5e-5), at 124 against 118 ms
    # per step
    f.set_online(250_000, epochs=1, batch=4096, lr=2e-4, above=0.5)
    # segments of the run: judged every 10 steps; when the trajectory is
    # compared at a step count that is not a multiple of 10, the run stops
    # there for the snapshot and the next segment ends on the next multiple
    cut = plain_end_step if plain_end is not None and 0 < plain_end_step <= steps else 0
    marks = sorted(set(list(range(10, steps + 1, 10)) + ([cut] if cut else [])))
    judged, mode, done, run_s, traj = [], LEARNED, 0, 0.0, None
    for mk in marks:
        n = mk - done
        ctx.synchronize()
        t1 = time.perf_counter()
        r = f.steps(dt, n, mode, judge_every=n if mk % 10 == 0 else 0, tol=1e-4)  # a fallback is for good
        ctx.synchronize()
        run_s += time.perf_counter() - t1
        done = mk
        mode = r["mode"]
        if mk % 10 == 0 and r["judged"] > 0:
            judged.append(r["last_error"])
            log(f"[bench] far field: steps {mk} judged error {r['last_error']:.3e} mode {mode}")
        if mk == cut:
            st = f.state()
            traj = {
                "step": mk,
                "vs": "plain FP32-compensated RK4 run from the same initial state",
                "position_max_abs": float(np.abs(np.stack([st["x"], st["y"], st["z"]]) -
                                                 np.stack([plain_end["x"], plain_end["y"], plain_end["z"]])).max()),
                "relvort_max_rel": float(np.abs(st["relvort"] - plain_end["relvort"]).max() /
                                         np.abs(plain_end["relvort"]).max()),
                "velocity_max_rel": rel((st["u"], st["v"], st["w"]),
                                        (plain_end["u"], plain_end["v"], plain_end["w"])),
            }
            log(f"[bench] far field: trajectory after {mk} steps vs the plain run: {traj}")
    ms_step = 1e3 * run_s / max(1, done)
    records = float(f.far_counts().mean())
    f.delete()
    return {
        "speedup_vs_plain": round(plain_ms_per_step / ms_step, 3),
        "max_rel_err": max(judged) if judged else None,
        "err_definition": "max_i |u_i - u_plain_i| / max_i |u_plain_i| at every judged step (plain = FP32 sum)",
        "ms_per_step": round(ms_step, 2),
        "plain_ms_per_step": round(plain_ms_per_step, 2),
        "steps": done,
        "judge_every": 10,
        "tol": 1e-4,
        "fell_back": mode != LEARNED,
        "judged_errors": judged,
        "includes": "judge sums, and online fine-tuning (250k samples, 1 epoch) at every judge whose error is "
                    "above half the tolerance",
        "fine_tuned_judges": int(sum(e > 0.5e-4 for e in judged)),
        "t0_error": {"without_network": e_exp, "with_network": e_net, "gain": round(e_exp / e_net, 3)},
        "trajectory_error": traj,
        "eval_ms": {"plain": round(ms_plain, 2), "expansion": round(ms_exp, 2), "learned": round(ms_net, 2)},
        "cells": int(f.ncell),
        "near_radius": near,
        "band_radius": band,
        "parent_radius": parent,
        "expansion_records_mean": records,
        "expansion_order": order,
        "near_cells_mean": float(nc.mean()),
        "band_cells_mean": float(bc.mean()),
        "training": {"samples": samples, "epochs": epochs, "seconds": round(train_s, 1), "loss": loss,
                     "states": "5 states 5 steps apart of an expansion-mode run, then back to t = 0",
                     "band_cell_draws": draws},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--full-error", type=int, default=1,
                    help="FP32 vs FP64 over every target at t = 0 and after the timed steps (0 = skip)")
    ap.add_argument("--fp64-steps", type=int, default=2, help="steps timed with FP64 pair sums (0 = skip)")
    ap.add_argument("--stream-steps", type=int, default=2,
                    help="steps timed with the stream functions in every step (0 = skip)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU baseline budget (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=default_cpu_threads(),
                    help="oracle threads (default: the CPUs this process may run on, capped by OMP_NUM_THREADS)")
    ap.add_argument("--graph", action="store_true", help="replay steps as hipGraphs (no per-kernel timing)")
    ap.add_argument("--workload", default="n655k",
                    help="n655k (default, the headline line) or cfg0/cfg1/cfg2/cfg3/pse_ops for the other "
                         "workloads; icos9_splits times the L9 sum at several source-split counts")
    ap.add_argument("--cfg-steps", type=int, default=0, help="override the step count of cfg1/cfg3")
    ap.add_argument("--ff-steps", type=int, default=100, help="learned far-field RK4 steps (0 = skip)")
    ap.add_argument("--ff-samples", type=int, default=10_000_000, help="far-field training samples")
    ap.add_argument("--ff-epochs", type=int, default=20)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 3 times (calibrates the HBM byte counters) and exit")
    args = ap.parse_args()

    if args.copy_calibration > 0:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "copies": 3}))
        return

    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; the direct sum would shard by target blocks "
                          "(each GPU all sources, a slice of the targets), --gpus N > 1 is not run"}))
        sys.exit(2)

    if args.workload != "n655k":
        import lpmgpu
        from lpmgpu import configs

        ctx = lpmgpu.Context(0)
        fn = {"cfg0": configs.run_cfg0, "cfg1": configs.run_cfg1, "cfg2": configs.run_cfg2,
              "cfg3": configs.run_cfg3, "pse_ops": configs.run_pse_ops,
              "icos9_splits": configs.run_icos9_splits}[args.workload]
        kw = {"steps": args.cfg_steps} if args.cfg_steps and args.workload in ("cfg1", "cfg3") else {}
        res = fn(ctx, log=log, **kw)
        if args.workload == "pse_ops" and args.cpu_seconds > 0:
            pse_ops_cpu_baseline(res, args.cpu_threads)
        elif args.cpu_seconds > 0 and args.workload != "icos9_splits":
            cfg_cpu_baseline(args.workload, res, args.cpu_seconds, args.cpu_threads)
        res["metric"] = METRIC
        print(json.dumps(res), flush=True)
        ctx.close()
        return

    import lpmgpu
    from lpmgpu import _lib as L

    t0 = time.perf_counter()
    m, zeta = build_workload(args.level)
    T, S = m.n, int(m.active.sum())
    pairs_per_sum = T * S - S
    log(f"[bench] mesh level {args.level}: T={T} S={S} built in {time.perf_counter() - t0:.1f}s")

    ctx = lpmgpu.Context(0)
    ctx.set_timing(not args.graph)
    omega = 2.0 * np.pi
    solver = lpmgpu.BVESolver(ctx, m.x, m.y, m.z, zeta, m.area, m.active, radius=1.0, rotation_rate=omega,
                              precision="fp32", use_graph=args.graph, project=True)
    log("[bench] solver ready (initial velocity computed)")

    # accuracy of the stored FP32 velocity against the FP64 GPU sum of the same
    # state, over every target, before the warm-up and after the timed steps
    err = {}
    if args.full_error:
        err["t0"] = velocity_error_fp64(ctx, solver.state(), m)
        log(f"[bench] FP32 vs FP64 over all {T} targets at t = 0: {err['t0']}")

    for _ in range(args.warmup):
        solver.timestep(args.dt)
    ctx.synchronize()
    ctx.reset_kernel_stats()
    log(f"[bench] warmup {args.warmup} done")

    t0 = time.perf_counter()
    for k in range(args.steps):
        solver.timestep(args.dt)
        log(f"[bench] step {k + 1}/{args.steps} done")
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    total_pairs = 4.0 * pairs_per_sum * args.steps
    value = total_pairs / elapsed / 1e9

    launches, kms = ctx.kernel_stats(L.K_SPHERE_BS)
    roofline = None
    if launches > 0:
        avg_s = kms / launches / 1e3
        achieved = FLOPS_PER_PAIR * pairs_per_sum / avg_s / 1e12
        roofline = {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": MEASURED_TRAFFIC_BYTES.get(args.level),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0)",
            "kernel": "pair_sum_kernel<SphereBS<float>>",
            "avg_launch_ms": round(avg_s * 1e3, 3),
            "launches": launches,
            "flops_per_launch": FLOPS_PER_PAIR * pairs_per_sum,
        }
    end_state = solver.state()
    if args.full_error:
        err["end"] = velocity_error_fp64(ctx, end_state, m)
        err["end_step"] = args.warmup + args.steps
        log(f"[bench] FP32 vs FP64 over all targets after {args.warmup + args.steps} steps: {err['end']}")
        err["vs"] = "FP64 GPU direct sum of the same state, every target"
    ke, ens, circ = solver.diagnostics()
    solver.delete()

    # the same steps at the reference's precision (FP64 pair sums), and with
    # the stream functions at the end of every step (SphereBVESolver.f90:352)
    other = {}
    for key, kw, nsteps in (("fp64", dict(precision="fp64"), args.fp64_steps),
                            ("with_stream", dict(precision="fp32", stream_each_step=True), args.stream_steps)):
        if nsteps <= 0:
            continue
        s2 = lpmgpu.BVESolver(ctx, m.x, m.y, m.z, zeta, m.area, m.active, absvort=zeta + 2 * omega * m.z,
                              radius=1.0, rotation_rate=omega, project=True, **kw)
        ctx.synchronize()
        t1 = time.perf_counter()
        s2.timestep(args.dt, nsteps)
        ctx.synchronize()
        other[key] = 1e3 * (time.perf_counter() - t1) / nsteps
        s2.delete()
        log(f"[bench] {key}: {other[key]:.1f} ms per step over {nsteps} steps")

    ff = None
    if args.ff_steps > 0:
        log("[bench] learned far field ...")
        ff = learned_far_field(ctx, m, zeta, omega, args.dt, ms_per_step, args.ff_steps, args.ff_samples,
                               args.ff_epochs, plain_end=end_state, plain_end_step=args.warmup + args.steps)
        log(f"[bench] learned far field: {ff['speedup_vs_plain']}x at max error {ff['max_rel_err']}")

    cpu = None
    if args.cpu_seconds > 0:
        log("[bench] CPU baseline ...")
        port = cpu_baseline(m, zeta, args.cpu_seconds, args.cpu_threads)
        log(f"[bench] CPU baseline (oracle) {port}")
        cpu = reference_cpu_baseline(m, zeta, args.cpu_seconds, args.cpu_threads)
        log(f"[bench] CPU baseline (reference) {cpu}")
        if cpu is None:
            cpu = port
        else:
            # the reference's own loop is the baseline; the oracle's restatement in C
            # (faster per core) is reported beside it
            cpu["cpu_model"] = port.pop("cpu_model")
            cpu["host"] = port.pop("host")
            cpu["port"] = port
            one = reference_cpu_baseline(m, zeta, min(3.0, args.cpu_seconds / 4), 1)
            if one is not None:
                cpu["one_core"] = {"value": one["value"], "sample": one["sample"]}
            log(f"[bench] CPU baseline (reference, one rank) {one}")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "G pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"sphere BVE plain direct-sum RK4, icosahedral level {args.level} "
                        f"(T={T}, S={S}), Gaussian vortex + RH4, FP32-compensated pairs, FP64 RK4",
            "targets": T,
            "sources": S,
            "pairs_per_step": 4 * pairs_per_sum,
            "dt": args.dt,
            "step": "one classical RK4 step: 4 Biot-Savart sums (stage 1 reuses the stored velocity) and FP64 "
                    "stage algebra with projected stages; the stream-function sum the reference ends each step "
                    "with (SphereBVESolver.f90:352) is not in it (see rk4_steps_per_s_with_stream)",
            "stream_functions_each_step": False,
        },
        "rk4_steps_per_s": round(1e3 / ms_per_step, 4),
        "rk4_steps_per_s_with_stream": round(1e3 / other["with_stream"], 4) if "with_stream" in other else None,
        "fp64_ms_per_step": round(other["fp64"], 1) if "fp64" in other else None,
        "fp64_rk4_steps_per_s": round(1e3 / other["fp64"], 4) if "fp64" in other else None,
        "velocity_error": err,
        "diagnostics": {"ke": ke, "enstrophy": ens, "circulation": circ},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "learned_far_field": ff,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
