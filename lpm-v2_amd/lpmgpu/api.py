"""Python host side over the C ABI: a context plus the reference's operators.

Names follow lpm-v2: `bve_sphere_velocity` (BVESphereVelocity /
SetVelocityOnMesh), `plane_velocity` (planarIncompressibleVelocity),
`pse_plane_laplacian` (PSEPlaneLaplacianAtParticles), `bve_sphere_stream`
(SetStreamFunctionsOnMesh), `plane_stream` (SetStreamFunctionOnMesh), and the
solvers `BVESolver` / `PlaneSolver` with New / Timestep / Delete semantics.
Target blocks are 0-based half-open [i0, i1); outputs are full-length arrays
whose entries outside the block are zero.
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _lib as L
from ._lib import LpmError, f64, i32, pd, pi, prec_code


class Context:
    """One HIP device, one stream, one workspace (lpm_create / lpm_destroy)."""

    def __init__(self, device: int = 0):
        self._lib = L.lib()
        h = ctypes.c_void_p()
        rc = self._lib.lpm_create(ctypes.byref(h), device)
        if rc != L.LPM_OK:
            raise LpmError(rc, self._lib.lpm_last_error(None).decode())
        self.handle = h
        self.device = device
        # objects created on this context (solvers, batches, networks): they
        # hold pointers into it and are deleted before the context is
        self._children = weakref.WeakSet()

    # -- error plumbing
    def check(self, rc: int):
        if rc != L.LPM_OK:
            raise LpmError(rc, self._lib.lpm_last_error(self.handle).decode())

    def adopt(self, obj):
        self._children.add(obj)
        return obj

    def close(self):
        if getattr(self, "handle", None):
            for child in list(self._children):
                child.delete()
            self._lib.lpm_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- sums
    def bve_sphere_velocity(self, x, y, z, relvort, area, active, radius=1.0, i0=0, i1=None,
                            precision="fp32", rotation_rate=0.0):
        x, y, z, q, a = map(f64, (x, y, z, relvort, area))
        act = i32(active)
        n = x.shape[0]
        i1 = n if i1 is None else i1
        u, v, w = np.zeros(n), np.zeros(n), np.zeros(n)
        self.check(self._lib.lpm_bve_sphere_velocity(self.handle, pd(u), pd(v), pd(w), pd(x), pd(y), pd(z),
                                                     pd(q), pd(a), radius, rotation_rate, pi(act), n, i0, i1,
                                                     prec_code(precision)))
        return u, v, w

    def plane_velocity(self, x, y, vort, area, active, i0=0, i1=None, precision="fp32"):
        x, y, q, a = map(f64, (x, y, vort, area))
        act = i32(active)
        n = x.shape[0]
        i1 = n if i1 is None else i1
        u, v = np.zeros(n), np.zeros(n)
        self.check(self._lib.lpm_plane_velocity(self.handle, pd(u), pd(v), pd(x), pd(y), pd(q), pd(a), pi(act),
                                                n, i0, i1, prec_code(precision)))
        return u, v

    def pse_plane_laplacian(self, x, y, f, area, active, eps, order=8, i0=0, i1=None, precision="fp32"):
        x, y, f, a = map(f64, (x, y, f, area))
        act = i32(active)
        n = x.shape[0]
        i1 = n if i1 is None else i1
        lap = np.zeros(n)
        self.check(self._lib.lpm_pse_plane_laplacian(self.handle, pd(lap), pd(x), pd(y), pd(f), pd(a), pi(act),
                                                     n, i0, i1, eps, order, prec_code(precision)))
        return lap

    def pse_sphere_laplacian(self, x, y, z, f, area, active, eps, radius=1.0, i0=0, i1=None, precision="fp32"):
        x, y, z, f, a = map(f64, (x, y, z, f, area))
        act = i32(active)
        n = x.shape[0]
        i1 = n if i1 is None else i1
        lap = np.zeros(n)
        self.check(self._lib.lpm_pse_sphere_laplacian(self.handle, pd(lap), pd(x), pd(y), pd(z), pd(f), pd(a),
                                                      pi(act), n, i0, i1, eps, radius, prec_code(precision)))
        return lap

    def pse_sphere_laplacian_at_points(self, x, y, z, f, area, active, eps, px, py, pz, pf, radius=1.0,
                                       precision="fp32"):
        x, y, z, f, a, px, py, pz, pf = map(f64, (x, y, z, f, area, px, py, pz, pf))
        act = i32(active)
        out = np.zeros(px.shape[0])
        self.check(self._lib.lpm_pse_sphere_laplacian_at_points(
            self.handle, pd(out), pd(px), pd(py), pd(pz), pd(pf), px.shape[0], pd(x), pd(y), pd(z), pd(f), pd(a),
            pi(act), x.shape[0], eps, radius, prec_code(precision)))
        return out

    def pse_plane_gradient(self, x, y, f, area, active, eps, i0=0, i1=None, precision="fp32"):
        x, y, f, a = map(f64, (x, y, f, area))
        act = i32(active)
        n = x.shape[0]
        i1 = n if i1 is None else i1
        gx, gy = np.zeros(n), np.zeros(n)
        self.check(self._lib.lpm_pse_plane_gradient(self.handle, pd(gx), pd(gy), pd(x), pd(y), pd(f), pd(a),
                                                    pi(act), n, i0, i1, eps, prec_code(precision)))
        return gx, gy

    def pse_sphere_gradient(self, x, y, z, f, area, active, eps, radius=1.0, i0=0, i1=None, precision="fp32"):
        x, y, z, f, a = map(f64, (x, y, z, f, area))
        act = i32(active)
        n = x.shape[0]
        i1 = n if i1 is None else i1
        g = [np.zeros(n) for _ in range(3)]
        self.check(self._lib.lpm_pse_sphere_gradient(self.handle, *map(pd, g), pd(x), pd(y), pd(z), pd(f), pd(a),
                                                     pi(act), n, i0, i1, eps, radius, prec_code(precision)))
        return tuple(g)

    def pse_plane_second_partials(self, x, y, gx, gy, area, active, eps, i0=0, i1=None, precision="fp32"):
        x, y, gx, gy, a = map(f64, (x, y, gx, gy, area))
        act = i32(active)
        n = x.shape[0]
        i1 = n if i1 is None else i1
        d = [np.zeros(n) for _ in range(3)]
        self.check(self._lib.lpm_pse_plane_second_partials(self.handle, *map(pd, d), pd(x), pd(y), pd(gx), pd(gy),
                                                           pd(a), pi(act), n, i0, i1, eps, prec_code(precision)))
        return tuple(d)

    def pse_plane_double_dot(self, x, y, u, v, area, active, eps, i0=0, i1=None, precision="fp32"):
        x, y, u, v, a = map(f64, (x, y, u, v, area))
        act = i32(active)
        n = x.shape[0]
        i1 = n if i1 is None else i1
        dd = np.zeros(n)
        self.check(self._lib.lpm_pse_plane_double_dot(self.handle, pd(dd), pd(x), pd(y), pd(u), pd(v), pd(a),
                                                      pi(act), n, i0, i1, eps, prec_code(precision)))
        return dd

    def pse_sphere_double_dot(self, x, y, z, u, v, w, area, active, eps, radius=1.0, i0=0, i1=None,
                              precision="fp32"):
        x, y, z, u, v, w, a = map(f64, (x, y, z, u, v, w, area))
        act = i32(active)
        n = x.shape[0]
        i1 = n if i1 is None else i1
        dd = np.zeros(n)
        self.check(self._lib.lpm_pse_sphere_double_dot(self.handle, pd(dd), pd(x), pd(y), pd(z), pd(u), pd(v),
                                                       pd(w), pd(a), pi(act), n, i0, i1, eps, radius,
                                                       prec_code(precision)))
        return dd

    def pse_sphere_divergence(self, x, y, z, u, v, w, area, active, eps, radius=1.0, i0=0, i1=None,
                              precision="fp32"):
        x, y, z, u, v, w, a = map(f64, (x, y, z, u, v, w, area))
        act = i32(active)
        n = x.shape[0]
        i1 = n if i1 is None else i1
        dv = np.zeros(n)
        self.check(self._lib.lpm_pse_sphere_divergence(self.handle, pd(dv), pd(x), pd(y), pd(z), pd(u), pd(v),
                                                       pd(w), pd(a), pi(act), n, i0, i1, eps, radius,
                                                       prec_code(precision)))
        return dv

    def pse_plane_interpolate(self, x, y, f, area, active, eps, px, py, precision="fp32"):
        x, y, f, a, px, py = map(f64, (x, y, f, area, px, py))
        act = i32(active)
        out = np.zeros(px.shape[0])
        self.check(self._lib.lpm_pse_plane_interpolate(self.handle, pd(out), pd(px), pd(py), px.shape[0], pd(x),
                                                       pd(y), pd(f), pd(a), pi(act), x.shape[0], eps,
                                                       prec_code(precision)))
        return out

    def pse_sphere_interpolate(self, x, y, z, f, area, active, eps, px, py, pz, radius=1.0, precision="fp32"):
        x, y, z, f, a, px, py, pz = map(f64, (x, y, z, f, area, px, py, pz))
        act = i32(active)
        out = np.zeros(px.shape[0])
        self.check(self._lib.lpm_pse_sphere_interpolate(self.handle, pd(out), pd(px), pd(py), pd(pz), px.shape[0],
                                                        pd(x), pd(y), pd(z), pd(f), pd(a), pi(act), x.shape[0], eps,
                                                        radius, prec_code(precision)))
        return out

    def swe_plane_rhs(self, x, y, vort, div, h, bottom, area, active, eps, i0=0, i1=None, precision="fp32"):
        """SWEPlaneRHSIntegrals: returns (u, v, double_dot, lap_surf); bottom = topoFn(x, y)."""
        x, y, q, d, h, b, a = map(f64, (x, y, vort, div, h, bottom, area))
        act = i32(active)
        n = x.shape[0]
        i1 = n if i1 is None else i1
        out = [np.zeros(n) for _ in range(4)]
        self.check(self._lib.lpm_swe_plane_rhs(self.handle, *map(pd, out), pd(x), pd(y), pd(q), pd(d), pd(h), pd(b),
                                               pd(a), pi(act), n, i0, i1, eps, prec_code(precision)))
        return tuple(out)

    def bve_sphere_stream(self, x, y, z, relvort, absvort, area, active, radius=1.0, i0=0, i1=None,
                          precision="fp32"):
        x, y, z, qr, qa, a = map(f64, (x, y, z, relvort, absvort, area))
        act = i32(active)
        n = x.shape[0]
        i1 = n if i1 is None else i1
        pr, pa = np.zeros(n), np.zeros(n)
        self.check(self._lib.lpm_bve_sphere_stream(self.handle, pd(pr), pd(pa), pd(x), pd(y), pd(z), pd(qr),
                                                   pd(qa), pd(a), radius, pi(act), n, i0, i1, prec_code(precision)))
        return pr, pa

    def plane_stream(self, x, y, vort, area, active, i0=0, i1=None, precision="fp32"):
        x, y, q, a = map(f64, (x, y, vort, area))
        act = i32(active)
        n = x.shape[0]
        i1 = n if i1 is None else i1
        psi = np.zeros(n)
        self.check(self._lib.lpm_plane_stream(self.handle, pd(psi), pd(x), pd(y), pd(q), pd(a), pi(act), n, i0,
                                              i1, prec_code(precision)))
        return psi

    def diagnostics(self, u, v, w, vort, area, active):
        u, v, q, a = map(f64, (u, v, vort, area))
        w = None if w is None else f64(w)
        act = i32(active)
        out = [ctypes.c_double() for _ in range(3)]
        self.check(self._lib.lpm_diagnostics(self.handle, pd(u), pd(v), pd(w), pd(q), pd(a), pi(act), u.shape[0],
                                             *[ctypes.byref(o) for o in out]))
        return tuple(o.value for o in out)

    # -- timing / tuning
    def set_timing(self, enable: bool):
        self.check(self._lib.lpm_set_timing(self.handle, int(bool(enable))))

    def kernel_stats(self, kind: int):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        self.check(self._lib.lpm_kernel_stats(self.handle, kind, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def reset_kernel_stats(self):
        self.check(self._lib.lpm_reset_kernel_stats(self.handle))

    def set_tuning(self, kind: int, precision, targets_per_lane=0, tile_sources=0, splits=0):
        self.check(self._lib.lpm_set_tuning(self.handle, kind, prec_code(precision), targets_per_lane,
                                            tile_sources, splits))

    def get_tuning(self, kind: int, precision):
        r, kt, s = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self.check(self._lib.lpm_get_tuning(self.handle, kind, prec_code(precision), ctypes.byref(r),
                                            ctypes.byref(kt), ctypes.byref(s)))
        return r.value, kt.value, s.value

    def synchronize(self):
        self.check(self._lib.lpm_synchronize(self.handle))


def load_balance(n_items: int, n_procs: int):
    """MPISetup LoadBalance (1-based inclusive index blocks)."""
    lib = L.lib()
    s = np.zeros(n_procs, np.int32)
    e = np.zeros(n_procs, np.int32)
    m = np.zeros(n_procs, np.int32)
    rc = lib.lpm_load_balance(n_items, n_procs, pi(s), pi(e), pi(m))
    if rc != L.LPM_OK:
        raise LpmError(rc, "lpm_load_balance: invalid arguments")
    return s, e, m


def pse_eps(max_edge_length: float, power: float = 0.75) -> float:
    lib = L.lib()
    out = ctypes.c_double()
    rc = lib.lpm_pse_eps(max_edge_length, power, ctypes.byref(out))
    if rc != L.LPM_OK:
        raise LpmError(rc, "lpm_pse_eps: invalid arguments")
    return out.value


class BVESolver:
    """Device-resident sphere BVE RK4 solver (BVESolver New / Timestep / Delete)."""

    def __init__(self, ctx: Context, x, y, z, relvort, area, active, absvort=None, u=None, v=None, w=None,
                 radius=1.0, rotation_rate=0.0, precision="fp32", stream_each_step=False, use_graph=False,
                 project=False):
        self.ctx = ctx
        arrs = [f64(a) for a in (x, y, z, relvort, area)]
        self._keep = arrs
        act = i32(active)
        absv = None if absvort is None else f64(absvort)
        uu = None if u is None else f64(u)
        vv = None if v is None else f64(v)
        ww = None if w is None else f64(w)
        self.n = arrs[0].shape[0]
        flags = (L.SOLVER_STREAM_EACH_STEP if stream_each_step else 0) | (L.SOLVER_USE_GRAPH if use_graph else 0)
        flags |= L.SOLVER_PROJECT_STAGES if project else 0
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.lpm_bve_solver_new(ctx.handle, ctypes.byref(h), self.n, pd(arrs[0]), pd(arrs[1]),
                                              pd(arrs[2]), pd(arrs[3]), pd(absv), pd(arrs[4]), pi(act), pd(uu),
                                              pd(vv), pd(ww), radius, rotation_rate, prec_code(precision), flags))
        self.handle = h
        ctx.adopt(self)

    def set_particles(self, x, y, z, relvort, area, active, absvort=None, u=None, v=None, w=None):
        """Remesh re-upload (lpm_bve_solver_set_particles): a new particle set of any size; the
        velocity is recomputed on the device unless u, v, w are given."""
        arrs = [f64(a) for a in (x, y, z, relvort, area)]
        act = i32(active)
        absv = None if absvort is None else f64(absvort)
        uvw = [None if a is None else f64(a) for a in (u, v, w)]
        n = arrs[0].shape[0]
        self.ctx.check(self.ctx._lib.lpm_bve_solver_set_particles(
            self.handle, n, pd(arrs[0]), pd(arrs[1]), pd(arrs[2]), pd(arrs[3]), pd(absv), pd(arrs[4]), pi(act),
            *[pd(a) for a in uvw]))
        self.n = n

    def timestep(self, dt: float, nsteps: int = 1):
        self.ctx.check(self.ctx._lib.lpm_bve_solver_steps(self.handle, dt, nsteps))

    def state(self):
        n = self.n
        out = [np.zeros(n) for _ in range(7)]
        self.ctx.check(self.ctx._lib.lpm_bve_solver_get_state(self.handle, *[pd(o) for o in out]))
        return dict(zip(["x", "y", "z", "relvort", "u", "v", "w"], out))

    def stream(self):
        pr, pa = np.zeros(self.n), np.zeros(self.n)
        self.ctx.check(self.ctx._lib.lpm_bve_solver_stream(self.handle, pd(pr), pd(pa)))
        return pr, pa

    def diagnostics(self):
        out = [ctypes.c_double() for _ in range(3)]
        self.ctx.check(self.ctx._lib.lpm_bve_solver_diagnostics(self.handle, *[ctypes.byref(o) for o in out]))
        return tuple(o.value for o in out)

    def delete(self):
        if getattr(self, "handle", None) and self.ctx.handle:
            self.ctx._lib.lpm_bve_solver_delete(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.delete()
        except Exception:
            pass


class PlaneSolver:
    """Device-resident planar incompressible RK4 solver (PlaneSolver New / Timestep / Delete)."""

    def __init__(self, ctx: Context, x, y, vort, area, active, u=None, v=None, precision="fp32",
                 stream_each_step=False, use_graph=False):
        self.ctx = ctx
        arrs = [f64(a) for a in (x, y, vort, area)]
        act = i32(active)
        uu = None if u is None else f64(u)
        vv = None if v is None else f64(v)
        self.n = arrs[0].shape[0]
        flags = (L.SOLVER_STREAM_EACH_STEP if stream_each_step else 0) | (L.SOLVER_USE_GRAPH if use_graph else 0)
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.lpm_plane_solver_new(ctx.handle, ctypes.byref(h), self.n, pd(arrs[0]), pd(arrs[1]),
                                                pd(arrs[2]), pd(arrs[3]), pi(act), pd(uu), pd(vv),
                                                prec_code(precision), flags))
        self.handle = h
        ctx.adopt(self)

    def set_particles(self, x, y, vort, area, active, u=None, v=None):
        """Remesh re-upload (lpm_plane_solver_set_particles)."""
        arrs = [f64(a) for a in (x, y, vort, area)]
        act = i32(active)
        uv = [None if a is None else f64(a) for a in (u, v)]
        n = arrs[0].shape[0]
        self.ctx.check(self.ctx._lib.lpm_plane_solver_set_particles(
            self.handle, n, pd(arrs[0]), pd(arrs[1]), pd(arrs[2]), pd(arrs[3]), pi(act), *[pd(a) for a in uv]))
        self.n = n

    def timestep(self, dt: float, nsteps: int = 1):
        self.ctx.check(self.ctx._lib.lpm_plane_solver_steps(self.handle, dt, nsteps))

    def state(self):
        out = [np.zeros(self.n) for _ in range(4)]
        self.ctx.check(self.ctx._lib.lpm_plane_solver_get_state(self.handle, *[pd(o) for o in out]))
        return dict(zip(["x", "y", "u", "v"], out))

    def stream(self):
        psi = np.zeros(self.n)
        self.ctx.check(self.ctx._lib.lpm_plane_solver_stream(self.handle, pd(psi)))
        return psi

    def diagnostics(self):
        out = [ctypes.c_double() for _ in range(3)]
        self.ctx.check(self.ctx._lib.lpm_plane_solver_diagnostics(self.handle, *[ctypes.byref(o) for o in out]))
        return tuple(o.value for o in out)

    def delete(self):
        if getattr(self, "handle", None) and self.ctx.handle:
            self.ctx._lib.lpm_plane_solver_delete(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.delete()
        except Exception:
            pass


class BVEBatch:
    """Many independent sphere BVE problems advanced together (lpm_bve_batch_*):
    one persistent launch per velocity evaluation, optional hipGraph per RK4 stage."""

    def __init__(self, ctx: Context, problems, precision="fp32", use_graph=False, project=False):
        """problems: list of dicts with x, y, z, relvort, area, active, radius, rotation_rate."""
        self.ctx = ctx
        self.sizes = [len(p["x"]) for p in problems]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        cat = {k: f64(np.concatenate([p[k] for p in problems])) for k in ("x", "y", "z", "relvort", "area")}
        act = i32(np.concatenate([p["active"] for p in problems]))
        rad = f64([p.get("radius", 1.0) for p in problems])
        om = f64([p.get("rotation_rate", 0.0) for p in problems])
        self.n = int(self.offsets[-1])
        flags = (L.SOLVER_USE_GRAPH if use_graph else 0) | (L.SOLVER_PROJECT_STAGES if project else 0)
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.lpm_bve_batch_new(ctx.handle, ctypes.byref(h), len(problems),
                                             self.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                             pd(cat["x"]), pd(cat["y"]), pd(cat["z"]), pd(cat["relvort"]),
                                             pd(cat["area"]), pi(act), pd(rad), pd(om), prec_code(precision), flags))
        self.handle = h
        ctx.adopt(self)

    def timestep(self, dt: float, nsteps: int = 1):
        self.ctx.check(self.ctx._lib.lpm_bve_batch_steps(self.handle, dt, nsteps))

    def state(self):
        out = [np.zeros(self.n) for _ in range(7)]
        self.ctx.check(self.ctx._lib.lpm_bve_batch_get_state(self.handle, *[pd(o) for o in out]))
        return dict(zip(["x", "y", "z", "relvort", "u", "v", "w"], out))

    def split(self, state):
        """Per-problem views of a concatenated state dict."""
        return [{k: v[a:b] for k, v in state.items()} for a, b in zip(self.offsets[:-1], self.offsets[1:])]

    def set_tuning(self, targets_per_lane, tile_sources, chunk_sources):
        self.ctx.check(self.ctx._lib.lpm_bve_batch_set_tuning(self.handle, targets_per_lane, tile_sources,
                                                              chunk_sources))

    def tuning(self):
        v = [ctypes.c_int32() for _ in range(4)]
        self.ctx.check(self.ctx._lib.lpm_bve_batch_get_tuning(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(zip(["targets_per_lane", "tile_sources", "chunk_sources", "n_items"], [x.value for x in v]))

    def autotune(self):
        v = [ctypes.c_int32() for _ in range(3)]
        ms = ctypes.c_double()
        self.ctx.check(self.ctx._lib.lpm_bve_batch_autotune(self.handle, *[ctypes.byref(x) for x in v],
                                                            ctypes.byref(ms)))
        return dict(targets_per_lane=v[0].value, tile_sources=v[1].value, chunk_sources=v[2].value, ms=ms.value)

    def pairs_per_eval(self) -> int:
        p = ctypes.c_int64()
        self.ctx.check(self.ctx._lib.lpm_bve_batch_pairs(self.handle, ctypes.byref(p)))
        return p.value

    def delete(self):
        if getattr(self, "handle", None) and self.ctx.handle:
            self.ctx._lib.lpm_bve_batch_delete(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.delete()
        except Exception:
            pass
