"""Ragged inputs: random particle counts (1 ... 2,500, most not multiples of a
wave), random active masks, random target blocks [i0, i1) and random point
sets, for the sphere and plane Biot-Savart sums, the PSE Laplacians, the
stream functions and the PSE derivative, interpolation and SWE sums, against
the FP64 oracle: FP64 to 1e-11, FP32 to 5e-4.  Random points come much
closer than mesh points, and FP32 position rounding (6e-8) over a separation
of 1e-3 moves a near pair's term by about 1e-4, so the FP32 bound is looser
than on meshes; it still fails on any missing or misplaced target.  Entries
outside the block stay zero, as the oracle leaves them."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEEDS = [11, 12, 13, 14, 15, 16]
SIZES = [1, 2, 63, 65, 257, 1000, 2500]
FP32_TOL = 5e-4
# FP32 PSE and SWE sums: outputs are differences of larger terms (the full-size
# tests hold them to 1e-4 on meshes); random points and small n add spread
PSE32_TOL = 2e-3


def _separated(p, dmin):
    """Drop each point closer than dmin to an earlier kept point.  The oracle
    evaluates the sphere denominator as R^2 - x.y, the reference's form, which
    for a pair at distance d keeps only about eps_64 / (d^2 / 2) of relative
    accuracy (1.5e-9 was seen at d ~ 1e-4); the library's |x - y|^2 / 2 form
    keeps full accuracy.  At d >= 0.01 the two agree to 1e-11."""
    keep = []
    for i in range(p.shape[1]):
        if not keep or np.min(np.sum((p[:, keep] - p[:, i:i + 1]) ** 2, axis=0)) >= dmin * dmin:
            keep.append(i)
    return p[:, keep]


def _sphere(rng, n):
    p = rng.normal(size=(3, n))
    p = _separated(p / np.linalg.norm(p, axis=0), 0.01)
    n = p.shape[1]
    area = rng.uniform(0.5, 1.5, n) * (4 * np.pi / max(n, 1))
    act = (rng.uniform(size=n) < 0.7).astype(np.int32)
    return p[0].copy(), p[1].copy(), p[2].copy(), area, act


def _plane(rng, n):
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    area = rng.uniform(0.5, 1.5, n) * (4.0 / max(n, 1))
    act = (rng.uniform(size=n) < 0.7).astype(np.int32)
    return x, y, area, act


def _block(rng, n):
    i0 = int(rng.integers(0, n))
    i1 = int(rng.integers(i0 + 1, n + 1))
    return i0, i1


def _close(got, ref, tol, what):
    got = [np.asarray(g) for g in (got if isinstance(got, tuple) else (got,))]
    ref = [np.asarray(r) for r in (ref if isinstance(ref, tuple) else (ref,))]
    scale = max(max(float(np.abs(r).max()) for r in ref), 1e-300)
    for k, (g, r) in enumerate(zip(got, ref)):
        err = float(np.abs(g - r).max()) / scale
        assert err <= tol, (what, k, err)


@pytest.mark.parametrize("seed", SEEDS)
def test_ragged_sums_match_oracle(ctx, oracle, seed):
    rng = np.random.default_rng(seed)
    for n_req in SIZES:
        x, y, z, area, act = _sphere(rng, n_req)
        n = x.size
        q = rng.normal(size=n)
        i0, i1 = _block(rng, n)
        tag = f"seed {seed} n {n} block [{i0}, {i1}) active {int(act.sum())}"
        ref = oracle.bve_sphere_velocity(x, y, z, q, area, act, 1.0, i0, i1)
        _close(ctx.bve_sphere_velocity(x, y, z, q, area, act, i0=i0, i1=i1, precision="fp64"), ref, 1e-11,
               "sphere fp64 " + tag)
        _close(ctx.bve_sphere_velocity(x, y, z, q, area, act, i0=i0, i1=i1, precision="fp32"), ref, FP32_TOL,
               "sphere fp32 " + tag)
        ref = oracle.bve_sphere_stream(x, y, z, q, q + z, area, act, 1.0, i0, i1)
        _close(ctx.bve_sphere_stream(x, y, z, q, q + z, area, act, i0=i0, i1=i1, precision="fp64"), ref, 1e-11,
               "sphere stream fp64 " + tag)
        _close(ctx.bve_sphere_stream(x, y, z, q, q + z, area, act, i0=i0, i1=i1, precision="fp32"), ref, FP32_TOL,
               "sphere stream fp32 " + tag)
        eps = 0.5 / max(n, 4) ** 0.25
        f = np.sin(2 * z) + x
        ref = oracle.pse_sphere_laplacian(x, y, z, f, area, act, eps, 1.0, i0, i1)
        _close(ctx.pse_sphere_laplacian(x, y, z, f, area, act, eps, i0=i0, i1=i1, precision="fp64"), ref, 1e-11,
               "sphere pse fp64 " + tag)
        ref = oracle.pse_sphere_gradient(x, y, z, f, area, act, eps, 1.0, i0, i1)
        _close(ctx.pse_sphere_gradient(x, y, z, f, area, act, eps, i0=i0, i1=i1, precision="fp64"), ref, 1e-11,
               "sphere gradient fp64 " + tag)

        px, py, parea, pact = _plane(rng, n)
        ref = oracle.plane_velocity(px, py, q, parea, pact, i0, i1)
        _close(ctx.plane_velocity(px, py, q, parea, pact, i0=i0, i1=i1, precision="fp64"), ref, 1e-11,
               "plane fp64 " + tag)
        _close(ctx.plane_velocity(px, py, q, parea, pact, i0=i0, i1=i1, precision="fp32"), ref, FP32_TOL,
               "plane fp32 " + tag)
        ref = oracle.plane_stream(px, py, q, parea, pact, i0, i1)
        _close(ctx.plane_stream(px, py, q, parea, pact, i0=i0, i1=i1, precision="fp64"), ref, 1e-11,
               "plane stream fp64 " + tag)
        pe = 2.0 / max(n, 4) ** 0.5
        g = np.exp(-2 * (px * px + py * py))
        for order in (8, 2):
            ref = oracle.pse_plane_laplacian(px, py, g, parea, pact, pe, order, i0, i1)
            _close(ctx.pse_plane_laplacian(px, py, g, parea, pact, pe, order=order, i0=i0, i1=i1,
                                           precision="fp64"), ref, 1e-11, f"plane pse{order} fp64 " + tag)
        ref = oracle.pse_plane_gradient(px, py, g, parea, pact, pe, i0, i1)
        _close(ctx.pse_plane_gradient(px, py, g, parea, pact, pe, i0=i0, i1=i1, precision="fp64"), ref, 1e-11,
               "plane gradient fp64 " + tag)
        hh = 1.0 + 0.1 * px
        ref = oracle.swe_plane_rhs(px, py, q, 0.3 * q, hh, 0.05 * py, parea, pact, pe, i0, i1)
        _close(ctx.swe_plane_rhs(px, py, q, 0.3 * q, hh, 0.05 * py, parea, pact, pe, i0=i0, i1=i1,
                                 precision="fp64"), ref, 1e-11, "swe fp64 " + tag)
        m = int(rng.integers(1, 300))
        qx, qy = rng.uniform(-1, 1, m), rng.uniform(-1, 1, m)
        ref = oracle.pse_plane_interpolate(px, py, g, parea, pact, pe, qx, qy)
        _close(ctx.pse_plane_interpolate(px, py, g, parea, pact, pe, qx, qy, precision="fp64"), ref, 1e-11,
               "plane interpolate fp64 " + tag)
        _close(ctx.pse_plane_interpolate(px, py, g, parea, pact, pe, qx, qy, precision="fp32"), ref, PSE32_TOL,
               "plane interpolate fp32 " + tag)


@pytest.mark.parametrize("seed", SEEDS[:3])
def test_ragged_fp32_pse_sums_match_oracle(ctx, oracle, seed):
    """The FP32 PSE, derivative and SWE kernels on ragged inputs: their tails
    run at 1 or 2 targets per lane, other shapes than the FP64 kernels."""
    rng = np.random.default_rng(100 + seed)
    for n_req in SIZES:
        x, y, z, area, act = _sphere(rng, n_req)
        n = x.size
        i0, i1 = _block(rng, n)
        tag = f"seed {seed} n {n} block [{i0}, {i1}) active {int(act.sum())}"
        eps = 0.5 / max(n, 4) ** 0.25
        f = np.sin(2 * z) + x
        ref = oracle.pse_sphere_laplacian(x, y, z, f, area, act, eps, 1.0, i0, i1)
        _close(ctx.pse_sphere_laplacian(x, y, z, f, area, act, eps, i0=i0, i1=i1, precision="fp32"), ref,
               PSE32_TOL, "sphere pse fp32 " + tag)
        ref = oracle.pse_sphere_gradient(x, y, z, f, area, act, eps, 1.0, i0, i1)
        _close(ctx.pse_sphere_gradient(x, y, z, f, area, act, eps, i0=i0, i1=i1, precision="fp32"), ref,
               PSE32_TOL, "sphere gradient fp32 " + tag)
        px, py, parea, pact = _plane(rng, n)
        q = rng.normal(size=n)
        pe = 2.0 / max(n, 4) ** 0.5
        g = np.exp(-2 * (px * px + py * py))
        ref = oracle.pse_plane_laplacian(px, py, g, parea, pact, pe, 8, i0, i1)
        _close(ctx.pse_plane_laplacian(px, py, g, parea, pact, pe, order=8, i0=i0, i1=i1, precision="fp32"), ref,
               PSE32_TOL, "plane pse8 fp32 " + tag)
        ref = oracle.pse_plane_gradient(px, py, g, parea, pact, pe, i0, i1)
        _close(ctx.pse_plane_gradient(px, py, g, parea, pact, pe, i0=i0, i1=i1, precision="fp32"), ref,
               PSE32_TOL, "plane gradient fp32 " + tag)
        hh = 1.0 + 0.1 * px
        ref = oracle.swe_plane_rhs(px, py, q, 0.3 * q, hh, 0.05 * py, parea, pact, pe, i0, i1)
        _close(ctx.swe_plane_rhs(px, py, q, 0.3 * q, hh, 0.05 * py, parea, pact, pe, i0=i0, i1=i1,
                                 precision="fp32"), ref, PSE32_TOL, "swe fp32 " + tag)
