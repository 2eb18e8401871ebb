"""The allocation-free sphere Biot-Savart entry point
(lpm_bve_sphere_velocity_noalloc) through the C ABI, against the allocating
entry point (lpm_bve_sphere_velocity) and the FP64 oracle.

Device arrays are allocated by the test (the entry point must not) with the
HIP runtime the library itself is linked against, reached through ctypes.
Two calls in a row, the second on a smaller particle set and a partial
target block, so the shared scratch array is reused with stale contents.
"""
import ctypes

import numpy as np
import pytest

from helpers import cfg0_inputs, rel_err_max
from lpmgpu import _lib as L

pytestmark = pytest.mark.gpu

# FP64: both entry points sum the same pairs in FP64 in a different split
# order; FP32: the workload's 1e-6 bar, against the oracle and each other.
TOL = {"fp64": 1e-12, "fp32": 1e-6}
ORACLE_TOL = {"fp64": 1e-11, "fp32": 1e-6}
H2D, D2H = 1, 2


def _hip():
    """The HIP runtime already loaded by liblpm_direct.so."""
    L.lib()
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64" in line:
                hip = ctypes.CDLL(line.split()[-1])
                hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
                hip.hipFree.argtypes = [ctypes.c_void_p]
                hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
                return hip
    raise RuntimeError("libamdhip64 not loaded")


class DeviceArrays:
    """hipMalloc'ed copies of host arrays, freed on exit."""

    def __init__(self, hip):
        self.hip = hip
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, n, dtype=np.float64):
        out = np.empty(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def __enter__(self):
        return self

    def __exit__(self, *a):
        for p in self.ptrs:
            self.hip.hipFree(p)


def _noalloc(ctx, m, zeta, i0, i1, prec, sentinel=7.0):
    with DeviceArrays(_hip()) as d:
        ins = [d.put(np.asarray(a, np.float64)) for a in (m.x, m.y, m.z, zeta, m.area)]
        act = d.put(np.asarray(m.active, np.int32))
        outs = [d.put(np.full(m.n, sentinel)) for _ in range(3)]
        rc = ctx._lib.lpm_bve_sphere_velocity_noalloc(ctx.handle, *outs, *ins, 1.0, 0.0, act, m.n, i0, i1,
                                                      L.prec_code(prec))
        ctx.check(rc)
        return [d.get(p, m.n) for p in outs]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_matches_allocating_entry_point(ctx, oracle, prec):
    # call 1: icosahedral L5 (30,722 particles), full block
    m1, z1 = cfg0_inputs(5)
    print(f"[noalloc] {prec} call 1: icos L5 n={m1.n}", flush=True)
    got1 = _noalloc(ctx, m1, z1, 0, m1.n, prec)
    ref1 = ctx.bve_sphere_velocity(m1.x, m1.y, m1.z, z1, m1.area, m1.active, precision=prec)
    e1 = rel_err_max(got1, ref1)
    print(f"[noalloc] {prec} L5 vs allocating entry: {e1:.3e}", flush=True)
    assert e1 < TOL[prec]
    # call 2: a smaller set (icosahedral L3, 1,922 particles) and a partial
    # block, so the scratch holds the previous call's larger contents
    m2, z2 = cfg0_inputs(3)
    i0, i1 = 101, 1500
    print(f"[noalloc] {prec} call 2: icos L3 n={m2.n} block [{i0}, {i1})", flush=True)
    got2 = _noalloc(ctx, m2, z2, i0, i1, prec)
    ref2 = ctx.bve_sphere_velocity(m2.x, m2.y, m2.z, z2, m2.area, m2.active, i0=i0, i1=i1, precision=prec)
    blk = slice(i0, i1)
    e2 = rel_err_max([g[blk] for g in got2], [r[blk] for r in ref2])
    orc = oracle.bve_sphere_velocity(m2.x, m2.y, m2.z, z2, m2.area, m2.active)
    e2o = rel_err_max([g[blk] for g in got2], [r[blk] for r in orc])
    print(f"[noalloc] {prec} L3 block vs allocating entry: {e2:.3e}, vs oracle: {e2o:.3e}", flush=True)
    assert e2 < TOL[prec]
    assert e2o < ORACLE_TOL[prec]
    # entries outside the block are untouched
    for g in got2:
        assert np.all(g[:i0] == 7.0) and np.all(g[i1:] == 7.0)


def test_noalloc_rejects_oversize_and_host_arrays(ctx):
    lib = ctx._lib
    n_big = 131073
    with DeviceArrays(_hip()) as d:
        dp = d.put(np.zeros(4))
        ap = d.put(np.ones(4, np.int32))
        rc = lib.lpm_bve_sphere_velocity_noalloc(ctx.handle, *[dp] * 8, 1.0, 0.0, ap, n_big, 0, n_big, 1)
        assert rc == L.LPM_ERR_ARG
        assert b"131072" in lib.lpm_last_error(ctx.handle)
    # host (numpy) arrays are refused before anything is launched
    h = np.zeros(4)
    ha = np.ones(4, np.int32)
    hp = ctypes.c_void_p(h.ctypes.data)
    rc = lib.lpm_bve_sphere_velocity_noalloc(ctx.handle, *[hp] * 8, 1.0, 0.0, ctypes.c_void_p(ha.ctypes.data),
                                             4, 0, 4, 1)
    assert rc == L.LPM_ERR_ARG
    assert b"not device memory" in lib.lpm_last_error(ctx.handle)
    # the context still works afterwards
    m, z = cfg0_inputs(2)
    u = ctx.bve_sphere_velocity(m.x, m.y, m.z, z, m.area, m.active, precision="fp64")
    assert np.all(np.isfinite(u))
