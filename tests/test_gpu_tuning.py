"""Tile shape and source-split overrides (lpm_set_tuning) change how a pair sum
is scheduled, never what it computes: every kind, at every targets-per-lane R,
compensated tile length KT and split count, agrees with its default schedule to
FP32 rounding (FP64: to 1e-12).  An R without a kernel instance fails with an
error instead of running a kernel on a grid sized for another R."""
import numpy as np
import pytest

from helpers import cfg0_inputs, rel_err_max
from lpmgpu import _lib as L
from lpmgpu import mesh
from lpmgpu.api import LpmError

pytestmark = pytest.mark.gpu

# (R, KT, splits); splits 0 = the library's choice
SHAPES = [(1, 64, 1), (2, 128, 3), (4, 256, 7), (8, 64, 2), (4, 512, 0)]


def _plane():
    x, y, area, h = mesh.grid_centres(48, 1.0, jitter=1e-3, seed=3)
    f = np.exp(-4.0 * (x * x + y * y)) + 0.3 * x
    act = np.ones(x.size, np.int32)
    return x, y, area, act, f, 2.0 * h


def _calls(ctx, prec):
    m, zeta = cfg0_inputs(4)
    x, y, area, act, f, eps = _plane()
    u, v = np.sin(2 * y) + x, np.cos(x) * y
    seps = float(m.max_edge_length ** 0.75)
    rng = np.random.default_rng(5)
    px, py = rng.uniform(-0.8, 0.8, 777), rng.uniform(-0.8, 0.8, 777)
    q = rng.normal(size=(3, 555))
    q /= np.linalg.norm(q, axis=0)
    return {
        "sphere_bs": (L.K_SPHERE_BS, lambda: ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active,
                                                                     precision=prec)),
        "plane_bs": (L.K_PLANE_BS, lambda: ctx.plane_velocity(x, y, f, area, act, precision=prec)),
        "pse8": (L.K_PSE8, lambda: ctx.pse_plane_laplacian(x, y, f, area, act, eps, order=8, precision=prec)),
        "pse2": (L.K_PSE2, lambda: ctx.pse_plane_laplacian(x, y, f, area, act, eps, order=2, precision=prec)),
        "sphere_green": (L.K_SPHERE_GREEN, lambda: ctx.bve_sphere_stream(m.x, m.y, m.z, zeta, zeta + m.z, m.area,
                                                                         m.active, precision=prec)),
        "plane_green": (L.K_PLANE_GREEN, lambda: ctx.plane_stream(x, y, f, area, act, precision=prec)),
        "pse_sphere8": (L.K_PSE_SPHERE8, lambda: ctx.pse_sphere_laplacian(m.x, m.y, m.z, zeta, m.area, m.active,
                                                                          seps, precision=prec)),
        "plane_gradient": (L.K_PSE_DERIV_PLANE, lambda: ctx.pse_plane_gradient(x, y, f, area, act, eps,
                                                                               precision=prec)),
        "plane_double_dot": (L.K_PSE_DERIV_PLANE, lambda: ctx.pse_plane_double_dot(x, y, u, v, area, act, eps,
                                                                                   precision=prec)),
        "sphere_gradient": (L.K_PSE_DERIV_SPHERE, lambda: ctx.pse_sphere_gradient(m.x, m.y, m.z, zeta, m.area,
                                                                                  m.active, seps, precision=prec)),
        "sphere_divergence": (L.K_PSE_DERIV_SPHERE,
                              lambda: ctx.pse_sphere_divergence(m.x, m.y, m.z, m.y, -m.x, 0.2 * m.z, m.area,
                                                                m.active, seps, precision=prec)),
        "plane_interpolate": (L.K_PSE_INTERP_PLANE, lambda: ctx.pse_plane_interpolate(x, y, f, area, act, eps, px,
                                                                                      py, precision=prec)),
        "sphere_interpolate": (L.K_PSE_INTERP_SPHERE,
                               lambda: ctx.pse_sphere_interpolate(m.x, m.y, m.z, zeta, m.area, m.active, seps, *q,
                                                                  precision=prec)),
        "swe": (L.K_SWE_PLANE, lambda: ctx.swe_plane_rhs(x, y, f, 0.1 * f, 1.0 + 0.1 * x, 0.05 * y, area, act, eps,
                                                         precision=prec)),
    }


def _as_rows(out):
    if isinstance(out, np.ndarray):
        return [out]
    return [np.asarray(o) for o in out]


OPS_KINDS = (L.K_PSE_DERIV_PLANE, L.K_PSE_DERIV_SPHERE, L.K_PSE_INTERP_PLANE, L.K_PSE_INTERP_SPHERE, L.K_SWE_PLANE)


# FP32: the Biot-Savart and Green sums to 2e-6; the PSE and SWE sums, whose
# outputs are differences of larger terms, to 5e-5 (their FP32-vs-FP64 error is
# below 3e-5 at full size)
LOOSE = ("pse8", "pse2", "pse_sphere8", "plane_gradient", "plane_double_dot", "sphere_gradient",
         "sphere_divergence", "plane_interpolate", "sphere_interpolate", "swe")


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_tile_shapes_and_splits_agree(ctx, prec):
    for name, (kind, call) in _calls(ctx, prec).items():
        tol = 1e-12 if prec == "fp64" else (5e-5 if name in LOOSE else 2e-6)
        ref = _as_rows(call())
        try:
            for r, kt, sp in SHAPES:
                ctx.set_tuning(kind, prec, r, kt, sp)
                if prec == "fp64" and r == 8 and kind in OPS_KINDS:
                    with pytest.raises(LpmError):
                        call()
                    continue
                got = _as_rows(call())
                for k, (g, e) in enumerate(zip(got, ref)):
                    err = rel_err_max(g, e)
                    assert err <= tol, (name, r, kt, sp, k, err)
        finally:
            ctx.set_tuning(kind, prec, 0, 0, 0)
