"""ctypes loader for the FP64 CPU oracle (liblpm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product library.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblpm_oracle.so")

_d = ctypes.c_double
_i = ctypes.c_int
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)

_SIGS = {
    "lpm_oracle_set_mutation": ([_i], _i),
    "lpm_oracle_set_threads": ([_i], _i),
    "lpm_oracle_bve_sphere_velocity": ([_i, _pd, _pd, _pd, _pd, _pd, _pi, _d, _i, _i, _pd, _pd, _pd], None),
    "lpm_oracle_plane_velocity": ([_i, _pd, _pd, _pd, _pd, _pi, _i, _i, _pd, _pd], None),
    "lpm_oracle_pse_plane_laplacian": ([_i, _pd, _pd, _pd, _pd, _pi, _d, _i, _i, _i, _pd], None),
    "lpm_oracle_pse_kernel": ([_i, _d], _d),
    "lpm_oracle_pse_sphere_laplacian": ([_i, _pd, _pd, _pd, _pd, _pd, _pi, _d, _d, _i, _i, _pd], None),
    "lpm_oracle_pse_delta_kernel": ([_d], _d),
    "lpm_oracle_pse_deriv_kernel": ([_d], _d),
    "lpm_oracle_pse_plane_gradient": ([_i, _pd, _pd, _pd, _pd, _pi, _d, _i, _i, _pd, _pd], None),
    "lpm_oracle_pse_sphere_gradient": ([_i, _pd, _pd, _pd, _pd, _pd, _pi, _d, _d, _i, _i, _pd, _pd, _pd], None),
    "lpm_oracle_pse_plane_second_partials": ([_i, _pd, _pd, _pd, _pd, _pd, _pi, _d, _i, _i, _pd, _pd, _pd], None),
    "lpm_oracle_pse_plane_double_dot": ([_i, _pd, _pd, _pd, _pd, _pd, _pi, _d, _i, _i, _pd], None),
    "lpm_oracle_pse_sphere_double_dot": ([_i, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _d, _d, _i, _i, _pd], None),
    "lpm_oracle_pse_sphere_divergence": ([_i, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _d, _d, _i, _i, _pd], None),
    "lpm_oracle_pse_plane_interpolate": ([_i, _pd, _pd, _pd, _pd, _pi, _d, _i, _pd, _pd, _pd], None),
    "lpm_oracle_pse_sphere_interpolate": ([_i, _pd, _pd, _pd, _pd, _pd, _pi, _d, _d, _i, _pd, _pd, _pd, _pd],
                                          None),
    "lpm_oracle_pse_sphere_laplacian_points": ([_i, _pd, _pd, _pd, _pd, _pd, _pi, _d, _d, _i, _pd, _pd, _pd, _pd,
                                                _pd], None),
    "lpm_oracle_swe_plane_rhs": ([_i, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _d, _i, _i, _pd, _pd, _pd, _pd],
                                 None),
    "lpm_oracle_bve_sphere_stream": ([_i, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _d, _i, _i, _pd, _pd], None),
    "lpm_oracle_plane_stream": ([_i, _pd, _pd, _pd, _pd, _pi, _i, _i, _pd], None),
    "lpm_oracle_bve_timestep": ([_i, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _d, _d, _d], None),
    "lpm_oracle_bve_timestep_ex": ([_i, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _d, _d, _d, _i], None),
    "lpm_oracle_plane_timestep": ([_i, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _d], None),
    "lpm_oracle_diagnostics": ([_i, _i, _pd, _pd, _pd, _pd, _pd, _pi, _pd], None),
    "lpm_oracle_load_balance": ([_i, _i, _pi, _pi, _pi], None),
    "lpm_oracle_pse_eps": ([_d, _d], _d),
    "lpm_oracle_cell_moments": ([_i, _pd, _pd, _pd, _pd, _pi, _i, _pd, _pd], None),
    "lpm_oracle_mlp_param_count": ([_i, _i, _i, _i], ctypes.c_long),
    "lpm_oracle_mlp_forward": ([_i, _i, _i, _i, _i, _pd, _i, _pd, _pd], None),
    "lpm_oracle_mlp_loss_grad": ([_i, _i, _i, _i, _i, _pd, _i, _pd, _pd, _pd], _d),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        for k, (a, r) in _SIGS.items():
            f = getattr(_lib, k)
            f.argtypes = a
            f.restype = r
    return _lib


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ii(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a):
    return a.ctypes.data_as(_pd)


def _q(a):
    return a.ctypes.data_as(_pi)


class mutation:
    """Context manager planting mistake `m` in the oracle."""

    def __init__(self, m):
        self.m = m

    def __enter__(self):
        self.old = lib().lpm_oracle_set_mutation(self.m)

    def __exit__(self, *a):
        lib().lpm_oracle_set_mutation(self.old)


def set_threads(n):
    return lib().lpm_oracle_set_threads(n)


def bve_sphere_velocity(x, y, z, relvort, area, active, radius=1.0, i0=0, i1=None):
    x, y, z, q, a = map(_f, (x, y, z, relvort, area))
    act = _ii(active)
    n = x.shape[0]
    i1 = n if i1 is None else i1
    u, v, w = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().lpm_oracle_bve_sphere_velocity(n, _p(x), _p(y), _p(z), _p(q), _p(a), _q(act), radius, i0, i1,
                                         _p(u), _p(v), _p(w))
    return u, v, w


def plane_velocity(x, y, vort, area, active, i0=0, i1=None):
    x, y, q, a = map(_f, (x, y, vort, area))
    act = _ii(active)
    n = x.shape[0]
    i1 = n if i1 is None else i1
    u, v = np.zeros(n), np.zeros(n)
    lib().lpm_oracle_plane_velocity(n, _p(x), _p(y), _p(q), _p(a), _q(act), i0, i1, _p(u), _p(v))
    return u, v


def pse_plane_laplacian(x, y, f, area, active, eps, order=8, i0=0, i1=None):
    x, y, f, a = map(_f, (x, y, f, area))
    act = _ii(active)
    n = x.shape[0]
    i1 = n if i1 is None else i1
    lap = np.zeros(n)
    lib().lpm_oracle_pse_plane_laplacian(n, _p(x), _p(y), _p(f), _p(a), _q(act), eps, order, i0, i1, _p(lap))
    return lap


def pse_kernel(order, rho):
    return lib().lpm_oracle_pse_kernel(order, rho)


def pse_sphere_laplacian(x, y, z, f, area, active, eps, radius=1.0, i0=0, i1=None):
    x, y, z, f, a = map(_f, (x, y, z, f, area))
    act = _ii(active)
    n = x.shape[0]
    i1 = n if i1 is None else i1
    lap = np.zeros(n)
    lib().lpm_oracle_pse_sphere_laplacian(n, _p(x), _p(y), _p(z), _p(f), _p(a), _q(act), eps, radius, i0, i1,
                                          _p(lap))
    return lap


def pse_delta_kernel(rho):
    return lib().lpm_oracle_pse_delta_kernel(rho)


def pse_deriv_kernel(rho):
    return lib().lpm_oracle_pse_deriv_kernel(rho)


def _block(x, i0, i1):
    n = x.shape[0]
    return n, (n if i1 is None else i1)


def pse_plane_gradient(x, y, f, area, active, eps, i0=0, i1=None):
    x, y, f, a = map(_f, (x, y, f, area))
    act = _ii(active)
    n, i1 = _block(x, i0, i1)
    gx, gy = np.zeros(n), np.zeros(n)
    lib().lpm_oracle_pse_plane_gradient(n, _p(x), _p(y), _p(f), _p(a), _q(act), eps, i0, i1, _p(gx), _p(gy))
    return gx, gy


def pse_sphere_gradient(x, y, z, f, area, active, eps, radius=1.0, i0=0, i1=None):
    x, y, z, f, a = map(_f, (x, y, z, f, area))
    act = _ii(active)
    n, i1 = _block(x, i0, i1)
    g = [np.zeros(n) for _ in range(3)]
    lib().lpm_oracle_pse_sphere_gradient(n, _p(x), _p(y), _p(z), _p(f), _p(a), _q(act), eps, radius, i0, i1,
                                         *map(_p, g))
    return tuple(g)


def pse_plane_second_partials(x, y, gx, gy, area, active, eps, i0=0, i1=None):
    x, y, gx, gy, a = map(_f, (x, y, gx, gy, area))
    act = _ii(active)
    n, i1 = _block(x, i0, i1)
    d = [np.zeros(n) for _ in range(3)]
    lib().lpm_oracle_pse_plane_second_partials(n, _p(x), _p(y), _p(gx), _p(gy), _p(a), _q(act), eps, i0, i1,
                                               *map(_p, d))
    return tuple(d)


def pse_plane_double_dot(x, y, u, v, area, active, eps, i0=0, i1=None):
    x, y, u, v, a = map(_f, (x, y, u, v, area))
    act = _ii(active)
    n, i1 = _block(x, i0, i1)
    dd = np.zeros(n)
    lib().lpm_oracle_pse_plane_double_dot(n, _p(x), _p(y), _p(u), _p(v), _p(a), _q(act), eps, i0, i1, _p(dd))
    return dd


def pse_sphere_double_dot(x, y, z, u, v, w, area, active, eps, radius=1.0, i0=0, i1=None):
    x, y, z, u, v, w, a = map(_f, (x, y, z, u, v, w, area))
    act = _ii(active)
    n, i1 = _block(x, i0, i1)
    dd = np.zeros(n)
    lib().lpm_oracle_pse_sphere_double_dot(n, _p(x), _p(y), _p(z), _p(u), _p(v), _p(w), _p(a), _q(act), eps,
                                           radius, i0, i1, _p(dd))
    return dd


def pse_sphere_divergence(x, y, z, u, v, w, area, active, eps, radius=1.0, i0=0, i1=None):
    x, y, z, u, v, w, a = map(_f, (x, y, z, u, v, w, area))
    act = _ii(active)
    n, i1 = _block(x, i0, i1)
    dv = np.zeros(n)
    lib().lpm_oracle_pse_sphere_divergence(n, _p(x), _p(y), _p(z), _p(u), _p(v), _p(w), _p(a), _q(act), eps,
                                           radius, i0, i1, _p(dv))
    return dv


def pse_plane_interpolate(x, y, f, area, active, eps, px, py):
    x, y, f, a, px, py = map(_f, (x, y, f, area, px, py))
    act = _ii(active)
    out = np.zeros(px.shape[0])
    lib().lpm_oracle_pse_plane_interpolate(x.shape[0], _p(x), _p(y), _p(f), _p(a), _q(act), eps, px.shape[0],
                                           _p(px), _p(py), _p(out))
    return out


def pse_sphere_interpolate(x, y, z, f, area, active, eps, px, py, pz, radius=1.0):
    x, y, z, f, a, px, py, pz = map(_f, (x, y, z, f, area, px, py, pz))
    act = _ii(active)
    out = np.zeros(px.shape[0])
    lib().lpm_oracle_pse_sphere_interpolate(x.shape[0], _p(x), _p(y), _p(z), _p(f), _p(a), _q(act), eps, radius,
                                            px.shape[0], _p(px), _p(py), _p(pz), _p(out))
    return out


def pse_sphere_laplacian_points(x, y, z, f, area, active, eps, px, py, pz, pf, radius=1.0):
    x, y, z, f, a, px, py, pz, pf = map(_f, (x, y, z, f, area, px, py, pz, pf))
    act = _ii(active)
    out = np.zeros(px.shape[0])
    lib().lpm_oracle_pse_sphere_laplacian_points(x.shape[0], _p(x), _p(y), _p(z), _p(f), _p(a), _q(act), eps,
                                                 radius, px.shape[0], _p(px), _p(py), _p(pz), _p(pf), _p(out))
    return out


def swe_plane_rhs(x, y, vort, div, h, bottom, area, active, eps, i0=0, i1=None):
    """Returns (u, v, double_dot, lap_surf)."""
    x, y, q, d, h, b, a = map(_f, (x, y, vort, div, h, bottom, area))
    act = _ii(active)
    n, i1 = _block(x, i0, i1)
    out = [np.zeros(n) for _ in range(4)]
    lib().lpm_oracle_swe_plane_rhs(n, _p(x), _p(y), _p(q), _p(d), _p(h), _p(b), _p(a), _q(act), eps, i0, i1,
                                   *map(_p, out))
    return tuple(out)


def bve_sphere_stream(x, y, z, relvort, absvort, area, active, radius=1.0, i0=0, i1=None):
    x, y, z, qr, qa, a = map(_f, (x, y, z, relvort, absvort, area))
    act = _ii(active)
    n = x.shape[0]
    i1 = n if i1 is None else i1
    pr, pa = np.zeros(n), np.zeros(n)
    lib().lpm_oracle_bve_sphere_stream(n, _p(x), _p(y), _p(z), _p(qr), _p(qa), _p(a), _q(act), radius, i0, i1,
                                       _p(pr), _p(pa))
    return pr, pa


def plane_stream(x, y, vort, area, active, i0=0, i1=None):
    x, y, q, a = map(_f, (x, y, vort, area))
    act = _ii(active)
    n = x.shape[0]
    i1 = n if i1 is None else i1
    psi = np.zeros(n)
    lib().lpm_oracle_plane_stream(n, _p(x), _p(y), _p(q), _p(a), _q(act), i0, i1, _p(psi))
    return psi


def bve_timestep(state: dict, area, active, radius, rotation_rate, dt, nsteps=1, project=False):
    """Advance a copy of state {x,y,z,relvort,u,v,w} by nsteps RK4 steps
    (project=True maps stage inputs and new positions onto the sphere)."""
    s = {k: _f(v).copy() for k, v in state.items()}
    a = _f(area)
    act = _ii(active)
    n = s["x"].shape[0]
    for _ in range(nsteps):
        lib().lpm_oracle_bve_timestep_ex(n, _p(s["x"]), _p(s["y"]), _p(s["z"]), _p(s["relvort"]), _p(s["u"]),
                                         _p(s["v"]), _p(s["w"]), _p(a), _q(act), radius, rotation_rate, dt,
                                         int(project))
    return s


def plane_timestep(state: dict, vort, area, active, dt, nsteps=1):
    s = {k: _f(v).copy() for k, v in state.items()}
    q = _f(vort)
    a = _f(area)
    act = _ii(active)
    n = s["x"].shape[0]
    for _ in range(nsteps):
        lib().lpm_oracle_plane_timestep(n, _p(s["x"]), _p(s["y"]), _p(q), _p(s["u"]), _p(s["v"]), _p(a), _q(act),
                                        dt)
    return s


def diagnostics(u, v, w, vort, area, active):
    u, v, q, a = map(_f, (u, v, vort, area))
    ncomp = 2 if w is None else 3
    w = np.zeros_like(u) if w is None else _f(w)
    act = _ii(active)
    out = np.zeros(3)
    lib().lpm_oracle_diagnostics(u.shape[0], ncomp, _p(u), _p(v), _p(w), _p(q), _p(a), _q(act), _p(out))
    return tuple(out)


def load_balance(n_items, n_procs):
    s = np.zeros(n_procs, np.int32)
    e = np.zeros(n_procs, np.int32)
    m = np.zeros(n_procs, np.int32)
    lib().lpm_oracle_load_balance(n_items, n_procs, _q(s), _q(e), _q(m))
    return s, e, m


def pse_eps(h, power=0.75):
    return lib().lpm_oracle_pse_eps(h, power)


def cell_moments(x, y, z, strength, cell, ncell, centroid):
    x, y, z, s = map(_f, (x, y, z, strength))
    c = _ii(cell)
    cen = _f(centroid).reshape(-1)
    m = np.zeros(10 * ncell)
    lib().lpm_oracle_cell_moments(x.shape[0], _p(x), _p(y), _p(z), _p(s), _q(c), ncell, _p(cen), _p(m))
    return m.reshape(ncell, 10)


def mlp_param_count(nin, width, nhidden, nout):
    return lib().lpm_oracle_mlp_param_count(nin, width, nhidden, nout)


def mlp_forward(params, inp, nin, width, nhidden, nout, act=0):
    p = _f(params)
    x = _f(inp).reshape(-1)
    nb = x.shape[0] // nin
    out = np.zeros(nb * nout)
    lib().lpm_oracle_mlp_forward(nin, width, nhidden, nout, act, _p(p), nb, _p(x), _p(out))
    return out.reshape(nb, nout)


def mlp_loss_grad(params, inp, target, nin, width, nhidden, nout, act=0):
    p = _f(params)
    x = _f(inp).reshape(-1)
    t = _f(target).reshape(-1)
    nb = x.shape[0] // nin
    g = np.zeros_like(p)
    loss = lib().lpm_oracle_mlp_loss_grad(nin, width, nhidden, nout, act, _p(p), nb, _p(x), _p(t), _p(g))
    return loss, g
