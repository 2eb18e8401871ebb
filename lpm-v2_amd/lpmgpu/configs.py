"""Measured runs of the workloads in lpmgpu.workloads (one JSON-able dict each).

Every run times the device work with the library's own HIP events (pair
kernel) and wall clock (whole steps, data resident in HBM), and states the
error of the FP32-compensated form against the plain FP64 direct sum on the
GPU, plus closed-form or analytic references where the workload has one.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib as L
from . import ic, workloads
from .api import BVEBatch, BVESolver, Context

PEAK_FP32_TFLOPS = 157.3
FLOPS = {"sphere": 23, "pse2": 11, "pse8": 18}


def _err(test, ref):
    t = np.atleast_2d(np.asarray(test, dtype=np.float64))
    r = np.atleast_2d(np.asarray(ref, dtype=np.float64))
    d = np.sqrt(((t - r) ** 2).sum(0))
    mag = np.sqrt((r**2).sum(0))
    return {
        "max_rel": float(d.max() / mag.max()),
        "rms_rel": float(np.sqrt((d**2).sum() / (mag**2).sum())),
        "pointwise_max_rel": float((d / np.maximum(mag, 1e-3 * mag.max())).max()),
    }


def _kernel(ctx, kind, flops_per_pair, pairs_per_launch):
    n, ms = ctx.kernel_stats(kind)
    if n == 0:
        return None
    avg = ms / n / 1e3
    tf = flops_per_pair * pairs_per_launch / avg / 1e12
    return {"launches": n, "avg_ms": round(avg * 1e3, 3), "tflops": round(tf, 2),
            "frac_fp32_peak": round(tf / PEAK_FP32_TFLOPS, 4)}


def run_sphere(ctx: Context, w: dict, precision: str, steps: int, project: bool, ref_precision="fp64",
               log=print):
    """Advance the workload with `precision` and with the FP64 reference side by
    side; report time/step, pair rate, and FP32-vs-FP64 errors at t=0 and at the end."""
    m = w["mesh"]
    T, S = m.n, m.n_active
    pairs = T * S - S
    ctx.set_timing(True)
    a = BVESolver(ctx, m.x, m.y, m.z, w["relvort"], m.area, m.active, rotation_rate=w["omega"],
                  precision=precision, project=project)
    s0 = a.state()
    out = {"targets": T, "sources": S, "pairs_per_sum": pairs, "dt": w["dt"], "steps": steps,
           "precision": precision, "projected_stages": project}
    ctx.reset_kernel_stats()
    t0 = time.perf_counter()
    a.timestep(w["dt"], steps)
    el = time.perf_counter() - t0
    out["ms_per_step"] = round(1e3 * el / steps, 3)
    out["rk4_steps_per_s"] = round(steps / el, 4)
    out["gpairs_per_s"] = round(4 * pairs * steps / el / 1e9, 2)
    out["pair_kernel"] = _kernel(ctx, L.K_SPHERE_BS, FLOPS["sphere"], pairs)
    s1 = a.state()
    out["diagnostics_end"] = dict(zip(["ke", "enstrophy", "circulation"], a.diagnostics()))
    a.delete()
    log(f"[{precision}] {out['ms_per_step']} ms/step")
    if ref_precision and ref_precision != precision:
        b = BVESolver(ctx, m.x, m.y, m.z, w["relvort"], m.area, m.active, rotation_rate=w["omega"],
                      precision=ref_precision, project=project)
        r0 = b.state()
        t0 = time.perf_counter()
        b.timestep(w["dt"], steps)
        out["ref_ms_per_step"] = round(1e3 * (time.perf_counter() - t0) / steps, 3)
        r1 = b.state()
        b.delete()
        out["velocity_error_t0"] = _err((s0["u"], s0["v"], s0["w"]), (r0["u"], r0["v"], r0["w"]))
        out["velocity_error_end"] = _err((s1["u"], s1["v"], s1["w"]), (r1["u"], r1["v"], r1["w"]))
        out["position_error_end"] = _err((s1["x"], s1["y"], s1["z"]), (r1["x"], r1["y"], r1["z"]))
        out["vs"] = f"plain {ref_precision} direct sum on the GPU, same steps"
    return out, s0, s1


def run_cfg0(ctx: Context, log=print):
    """cfg0: FP64 plain method at icosahedral L4 (also checked against the CPU oracle
    by the tests); 10 steps, dt 0.01, reference semantics (no projection)."""
    w = workloads.cfg0()
    out, s0, s1 = run_sphere(ctx, w, "fp64", w["steps"], project=False, ref_precision=None, log=log)
    out["workload"] = "cfg0 icosahedral L4 Gaussian vortex, FP64, dt 0.01 x 10"
    return out


def run_cfg1(ctx: Context, steps=None, log=print):
    """cfg1: cubed sphere L7 RH4, FP32-compensated pairs, FP64 RK4, dt 0.005 x 100,
    against the plain FP64 GPU sum; also the closed-form RH4 velocity at t = 0."""
    w = workloads.cfg1()
    n = steps or w["steps"]
    out, s0, s1 = run_sphere(ctx, w, "fp32", n, project=True, log=log)
    m = w["mesh"]
    ex = ic.rh4_velocity(m.x, m.y, m.z).T
    out["closed_form_error_t0"] = _err((s0["u"], s0["v"], s0["w"]), tuple(ex))
    out["workload"] = f"cfg1 cubed-sphere L7 RH4, FP32-compensated, dt 0.005 x {n}"
    return out


def run_cfg2(ctx: Context, log=print):
    """cfg2: 512^2 jittered grid PSE Laplacian, eps = 2h, FP32-compensated (orders 2 and 8)
    against the FP64 GPU sum and the analytic Laplacian."""
    w = workloads.cfg2()
    n = w["x"].shape[0]
    act = np.ones(n, np.int32)
    pairs = n * n
    res = {"workload": "cfg2 PSE Laplacian 512^2 jittered grid, eps = 2h", "particles": n, "pairs": pairs}
    r2 = w["x"] ** 2 + w["y"] ** 2
    inner = r2 < 0.5
    ctx.set_timing(True)
    for order, kind in ((2, L.K_PSE2), (8, L.K_PSE8)):
        ctx.reset_kernel_stats()
        t0 = time.perf_counter()
        g32 = ctx.pse_plane_laplacian(w["x"], w["y"], w["f"], w["area"], act, w["eps"], order=order, precision="fp32")
        t32 = time.perf_counter() - t0
        k = _kernel(ctx, kind, FLOPS[f"pse{order}"], pairs)
        t0 = time.perf_counter()
        g64 = ctx.pse_plane_laplacian(w["x"], w["y"], w["f"], w["area"], act, w["eps"], order=order, precision="fp64")
        t64 = time.perf_counter() - t0
        res[f"pse{order}"] = {
            "fp32_call_ms": round(1e3 * t32, 2),
            "fp64_call_ms": round(1e3 * t64, 2),
            "pair_kernel": k,
            "fp32_vs_fp64": _err(g32, g64),
            "fp64_vs_analytic_interior": _err(g64[inner], w["exact"][inner]),
            "fp32_vs_analytic_interior": _err(g32[inner], w["exact"][inner]),
        }
        log(f"[pse{order}] {res[f'pse{order}']}")
    # the same grid without jitter: the PSE consistency error against the analytic Laplacian
    w0 = workloads.cfg2(jitter=0.0)
    for order in (2, 8):
        g = ctx.pse_plane_laplacian(w0["x"], w0["y"], w0["f"], w0["area"], act, w0["eps"], order=order,
                                    precision="fp32")
        res[f"pse{order}_unjittered_fp32_vs_analytic_interior"] = _err(g[inner], w0["exact"][inner])
    return res


def run_cfg3(ctx: Context, steps=None, autotune=True, log=print):
    """cfg3: 64 uneven problems in one persistent launch per evaluation, a hipGraph
    per RK4 stage, tile shape from the autotuner, FP32-compensated; 20 steps."""
    w = workloads.cfg3()
    n = steps or w["steps"]
    ctx.set_timing(False)
    b = BVEBatch(ctx, w["problems"], precision="fp32", use_graph=True, project=True)
    ppe = b.pairs_per_eval()
    out = {"workload": f"cfg3 64 icosahedral problems L3-L7 (seeded), FP32-compensated, dt 0.01 x {n}",
           "problems": len(w["problems"]), "levels": [p["level"] for p in w["problems"]],
           "pairs_per_eval": ppe}
    if autotune:
        t0 = time.perf_counter()
        out["autotune"] = b.autotune()
        out["autotune_s"] = round(time.perf_counter() - t0, 2)
    out["tuning"] = b.tuning()
    b.timestep(w["dt"], 1)  # captures the stage graphs
    t0 = time.perf_counter()
    b.timestep(w["dt"], n)
    el = time.perf_counter() - t0
    out["ms_per_step"] = round(1e3 * el / n, 3)
    out["gpairs_per_s"] = round(4 * ppe * n / el / 1e9, 2)
    out["frac_fp32_peak_step"] = round(23 * 4 * ppe * n / el / 1e12 / PEAK_FP32_TFLOPS, 4)
    st = b.split(b.state())
    b.delete()
    # accuracy: the same batch in FP64 for 1 step from the initial state vs FP32
    b32 = BVEBatch(ctx, w["problems"], precision="fp32", project=True)
    b64 = BVEBatch(ctx, w["problems"], precision="fp64", project=True)
    s32, s64 = b32.state(), b64.state()
    out["velocity_error_t0_vs_fp64"] = _err((s32["u"], s32["v"], s32["w"]), (s64["u"], s64["v"], s64["w"]))
    b32.delete()
    b64.delete()
    out["finite_after_steps"] = bool(all(np.isfinite(s["u"]).all() for s in st))
    return out
