"""PSE gradient / second partials / double dot / divergence / interpolation and
the planar SWE right-hand side on the GPU, through the C ABI, against the
FP64 oracle (PSEDirectSum.f90:121-418, 537-579; SWEPlaneSolver.f90:457-538)
and the reference's spherePSETest thresholds.

Tolerances (max abs difference over the max abs oracle value): FP64 1e-11
(summation order and the split (c_j + c_i) = c_j + c_i of the sums differ);
FP32 with compensated accumulation 2e-4 for derivative sums (FP32
separations of points 1/eps apart in kernel units), 2e-5 for interpolation.
"""
import numpy as np
import pytest

import lpmgpu
from helpers import icos
from lpmgpu import mesh
from test_oracle_pins_pse_ops import harmonic54, harmonic54_gradient, latlon_grid

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-11, "fp32": 2e-4}


def _rel(got, ref):
    got = np.atleast_2d(np.asarray(got))
    ref = np.atleast_2d(np.asarray(ref))
    return float(np.abs(got - ref).max() / np.abs(ref).max())


def _plane(n_side=64, jitter=0.25, seed=2):
    x, y, area, h = mesh.grid_centres(n_side, 1.0, jitter=jitter * 2.0 / n_side, seed=seed)
    act = np.ones(x.size, np.int32)
    act[::7] = 0  # some passive particles (targets, not sources)
    return x, y, area, act, 2.5 * h


def _sphere(level=4):
    m = icos(level)
    return m, harmonic54(m.x, m.y, m.z), lpmgpu.pse_eps(m.max_edge_length, 0.6)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_plane_gradient(ctx, oracle, precision):
    x, y, area, act, eps = _plane()
    f = np.sin(3 * x) * np.cos(2 * y) + 0.5 * x
    ref = oracle.pse_plane_gradient(x, y, f, area, act, eps)
    got = ctx.pse_plane_gradient(x, y, f, area, act, eps, precision=precision)
    e = _rel(got, ref)
    print(precision, "plane gradient", e)
    assert e < TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_plane_second_partials(ctx, oracle, precision):
    x, y, area, act, eps = _plane()
    gx, gy = np.cos(2 * x) * y, x * x - np.sin(y)
    ref = oracle.pse_plane_second_partials(x, y, gx, gy, area, act, eps)
    got = ctx.pse_plane_second_partials(x, y, gx, gy, area, act, eps, precision=precision)
    e = _rel(got, ref)
    print(precision, "second partials", e)
    assert e < TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_plane_double_dot(ctx, oracle, precision):
    x, y, area, act, eps = _plane()
    u, v = np.sin(2 * y) + x, np.cos(x) * y
    ref = oracle.pse_plane_double_dot(x, y, u, v, area, act, eps)
    got = ctx.pse_plane_double_dot(x, y, u, v, area, act, eps, precision=precision)
    e = _rel(got, ref)
    print(precision, "plane double dot", e)
    assert e < TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("level,power", [(4, 0.6), (5, 0.9)])
def test_sphere_gradient(ctx, oracle, precision, level, power):
    """power 0.9 at nest 5 gives 7 eps <= 0.56 R: the FP32 series distance."""
    m = icos(level)
    f = harmonic54(m.x, m.y, m.z)
    eps = lpmgpu.pse_eps(m.max_edge_length, power)
    i1 = min(m.n, 6000)
    ref = oracle.pse_sphere_gradient(m.x, m.y, m.z, f, m.area, m.active, eps, i1=i1)
    got = ctx.pse_sphere_gradient(m.x, m.y, m.z, f, m.area, m.active, eps, i1=i1, precision=precision)
    e = _rel([g[:i1] for g in got], [r[:i1] for r in ref])
    print(precision, level, "sphere gradient", e, "eps", eps)
    assert e < TOL[precision]


def test_fp64_series_distance_matches_oracle(ctx, oracle):
    """eps = 0.045 at nest 5 gives 7 eps <= 0.35 R: the FP64 kernels take the
    10-term series of the great-circle distance, the oracle asin."""
    m = icos(5)
    f = harmonic54(m.x, m.y, m.z)
    eps = 0.045
    u, v, w = harmonic54_gradient(m.x, m.y, m.z)
    i1 = 5000
    cases = {
        "laplacian": (oracle.pse_sphere_laplacian(m.x, m.y, m.z, f, m.area, m.active, eps, 1.0, 0, i1),
                      ctx.pse_sphere_laplacian(m.x, m.y, m.z, f, m.area, m.active, eps, i1=i1, precision="fp64")),
        "gradient": (oracle.pse_sphere_gradient(m.x, m.y, m.z, f, m.area, m.active, eps, i1=i1),
                     ctx.pse_sphere_gradient(m.x, m.y, m.z, f, m.area, m.active, eps, i1=i1, precision="fp64")),
        "divergence": (oracle.pse_sphere_divergence(m.x, m.y, m.z, u, v, w, m.area, m.active, eps, i1=i1),
                       ctx.pse_sphere_divergence(m.x, m.y, m.z, u, v, w, m.area, m.active, eps, i1=i1,
                                                 precision="fp64")),
    }
    qx, qy, qz = latlon_grid()
    sel = slice(0, None, 7)
    cases["interpolation"] = (
        oracle.pse_sphere_interpolate(m.x, m.y, m.z, f, m.area, m.active, eps, qx[sel], qy[sel], qz[sel]),
        ctx.pse_sphere_interpolate(m.x, m.y, m.z, f, m.area, m.active, eps, qx[sel], qy[sel], qz[sel],
                                   precision="fp64"))
    for name, (ref, got) in cases.items():
        ref = np.atleast_2d(np.asarray(ref))[:, :i1] if name != "interpolation" else ref
        got = np.atleast_2d(np.asarray(got))[:, :i1] if name != "interpolation" else got
        e = _rel(got, ref)
        print("fp64 series", name, e)
        assert e < 1e-11, (name, e)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sphere_double_dot_and_divergence(ctx, oracle, precision):
    m, f, eps = _sphere(4)
    u, v, w = harmonic54_gradient(m.x, m.y, m.z)
    w = w + 0.3 * m.z  # w differs from v, so the reference's w row is exercised
    ref = oracle.pse_sphere_double_dot(m.x, m.y, m.z, u, v, w, m.area, m.active, eps)
    got = ctx.pse_sphere_double_dot(m.x, m.y, m.z, u, v, w, m.area, m.active, eps, precision=precision)
    e1 = _rel(got, ref)
    ref = oracle.pse_sphere_divergence(m.x, m.y, m.z, u, v, w, m.area, m.active, eps)
    got = ctx.pse_sphere_divergence(m.x, m.y, m.z, u, v, w, m.area, m.active, eps, precision=precision)
    e2 = _rel(got, ref)
    print(precision, "sphere double dot", e1, "divergence", e2)
    assert e1 < TOL[precision] and e2 < TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_interpolation_at_points(ctx, oracle, precision):
    x, y, area, act, eps = _plane()
    f = np.exp(-3 * (x**2 + y**2))
    rng = np.random.default_rng(9)
    px, py = rng.uniform(-1.2, 1.2, 1000), rng.uniform(-1.2, 1.2, 1000)
    ref = oracle.pse_plane_interpolate(x, y, f, area, act, eps, px, py)
    got = ctx.pse_plane_interpolate(x, y, f, area, act, eps, px, py, precision=precision)
    e1 = _rel(got, ref)
    m, fs, eps = _sphere(4)
    qx, qy, qz = latlon_grid()
    sel = slice(0, None, 13)
    ref = oracle.pse_sphere_interpolate(m.x, m.y, m.z, fs, m.area, m.active, eps, qx[sel], qy[sel], qz[sel])
    got = ctx.pse_sphere_interpolate(m.x, m.y, m.z, fs, m.area, m.active, eps, qx[sel], qy[sel], qz[sel],
                                     precision=precision)
    e2 = _rel(got, ref)
    print(precision, "interpolation plane", e1, "sphere", e2)
    tol = 1e-11 if precision == "fp64" else 2e-5
    assert e1 < tol and e2 < tol


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_reference_interpolation_and_grid_laplacian_thresholds(ctx, precision):
    """spherePSETest (tests/CMakeLists.txt:74-75): icosahedral nest 3, eps = h^0.6;
    particles L_inf <= 0.0176 and squared L2 <= 2e-4 over rank 0's block of 4,
    lat-lon grid interpolation L_inf <= 0.018 and Laplacian L_inf <= 0.01805
    (SpherePSEConvTest.f90:383-389)."""
    m = icos(3)
    f = harmonic54(m.x, m.y, m.z)
    eps = lpmgpu.pse_eps(m.max_edge_length, 0.6)
    interp = ctx.pse_sphere_interpolate(m.x, m.y, m.z, f, m.area, m.active, eps, m.x, m.y, m.z,
                                        precision=precision)
    err = np.abs(interp - f)
    linf = float(err.max() / np.abs(f).max())
    s, e, _ = lpmgpu.load_balance(m.n, 4)
    blk = np.zeros(m.n, bool)
    blk[s[0] - 1: e[0]] = True
    a = (m.active == 1) & blk
    l2 = float((err[a] ** 2 * m.area[a]).sum() / (f[a] ** 2 * m.area[a]).sum())
    px, py, pz = latlon_grid()
    fg = harmonic54(px, py, pz)
    gi = ctx.pse_sphere_interpolate(m.x, m.y, m.z, f, m.area, m.active, eps, px, py, pz, precision=precision)
    ulinf = float(np.abs(gi - fg).max() / np.abs(fg).max())
    gl = ctx.pse_sphere_laplacian_at_points(m.x, m.y, m.z, f, m.area, m.active, eps, px, py, pz, fg,
                                            precision=precision)
    ulap = float(np.abs(gl + 30 * fg).max() / np.abs(30 * fg).max())
    print(precision, "linf", linf, "l2", l2, "grid interp", ulinf, "grid lap", ulap)
    assert linf <= 0.0176 and l2 <= 2e-4 and ulinf <= 0.018 and ulap <= 0.01805


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_swe_plane_rhs(ctx, oracle, precision):
    x, y, area, act, eps = _plane(48)
    vort = 3.0 * np.exp(-4 * (x**2 + y**2))
    div = 0.5 * np.sin(2 * x) * np.cos(y)
    h = 1.0 + 0.1 * np.exp(-6 * ((x - 0.2) ** 2 + y**2))
    bottom = 0.05 * x - 0.02 * y * y
    ref = oracle.swe_plane_rhs(x, y, vort, div, h, bottom, area, act, eps)
    got = ctx.swe_plane_rhs(x, y, vort, div, h, bottom, area, act, eps, precision=precision)
    errs = [_rel(g, r) for g, r in zip(got, ref)]
    print(precision, "swe u v ddot lap", errs)
    tol = 1e-11 if precision == "fp64" else 1e-4
    assert max(errs) < tol


def test_blocks_concatenate(ctx):
    x, y, area, act, eps = _plane(32)
    f = np.sin(3 * x) + y
    full = ctx.pse_plane_gradient(x, y, f, area, act, eps, precision="fp64")
    parts = [np.zeros(x.size), np.zeros(x.size)]
    for i0, i1 in ((0, 100), (100, 700), (700, x.size)):
        got = ctx.pse_plane_gradient(x, y, f, area, act, eps, i0=i0, i1=i1, precision="fp64")
        for k in range(2):
            assert np.all(got[k][:i0] == 0.0) and np.all(got[k][i1:] == 0.0)
            parts[k][i0:i1] = got[k][i0:i1]
    assert _rel(parts, full) < 1e-13


def test_empty_and_bad_arguments(ctx):
    x, y, area, act, eps = _plane(16)
    f = x + y
    assert ctx.pse_plane_interpolate(x, y, f, area, act, eps, np.zeros(0), np.zeros(0)).shape == (0,)
    gx, gy = ctx.pse_plane_gradient(x, y, f, area, act, eps, i0=5, i1=5)
    assert not gx.any() and not gy.any()
    with pytest.raises(lpmgpu.LpmError):
        ctx.pse_plane_gradient(x, y, f, area, act, -1.0)
    with pytest.raises(lpmgpu.LpmError):
        ctx.pse_plane_double_dot(x, y, f, f, area, act, eps, i0=3, i1=x.size + 1)
    with pytest.raises(lpmgpu.LpmError):
        ctx.swe_plane_rhs(x, y, f, f, f, f, area, act, eps, precision=7)
    # no active sources: every sum is zero
    none = np.zeros(x.size, np.int32)
    u, v, dd, lap = ctx.swe_plane_rhs(x, y, f, f, f, f, area, none, eps)
    assert not u.any() and not dd.any() and not lap.any()
    # the context survives
    ctx.pse_plane_gradient(x, y, f, area, act, eps)
