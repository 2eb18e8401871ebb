"""GPU parity: the HIP library (through the C ABI) against the FP64 CPU oracle.

Sizes are ones the oracle finishes in well under a second.  Tolerances:
FP64 pair kernels agree with the oracle to rounding (the GPU sums in a
different order and evaluates R^2 - x.y as |x - y|^2/2 + (R^2-|x|^2)/2 +
(R^2-|y|^2)/2); FP32 pair kernels with compensated accumulation must stay
within 1e-6 normwise relative error (velocity) as the workload states.
"""
import numpy as np
import pytest

from helpers import cfg0_inputs, icos, rel_err_max, rel_err_rms
from lpmgpu import LpmError, ic, load_balance, mesh

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11
FP32_VEL_TOL = 1e-6


def _print(name, **kw):
    print(name, " ".join(f"{k}={v:.3e}" if isinstance(v, float) else f"{k}={v}" for k, v in kw.items()), flush=True)


# ---------------------------------------------------------------- sphere BS
def test_sphere_velocity_fp64_icos3(ctx, oracle):
    m = icos(3)
    rng = np.random.default_rng(1)
    zeta = rng.normal(size=m.n)
    ref = oracle.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active)
    got = ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision="fp64")
    e = rel_err_max(got, ref)
    _print("sphere fp64 icos3", err=e)
    assert e < FP64_TOL


def test_sphere_velocity_cfg0_both_precisions(ctx, oracle):
    m, zeta = cfg0_inputs(4)
    ref = oracle.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active)
    g64 = ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision="fp64")
    g32 = ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision="fp32")
    e64, e32 = rel_err_max(g64, ref), rel_err_max(g32, ref)
    _print("sphere cfg0", fp64=e64, fp32=e32, fp32_rms=rel_err_rms(g32, ref))
    assert e64 < FP64_TOL
    assert e32 < FP32_VEL_TOL


def test_sphere_velocity_off_sphere_stage_points(ctx, oracle):
    """RK stage points x + dt u / 2 leave the sphere (SphereBVESolver.f90:249-329);
    the GPU's |x-y|^2/2 + (R^2-|x|^2)/2 + (R^2-|y|^2)/2 denominator must match
    the reference's R^2 - x.y there."""
    m, zeta = cfg0_inputs(4)
    u = np.array(oracle.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active))
    x, y, z = (np.stack([m.x, m.y, m.z]) + 0.5 * 0.01 * u)
    assert np.abs(np.sqrt(x * x + y * y + z * z) - 1).max() > 1e-6  # really off the sphere
    ref = oracle.bve_sphere_velocity(x, y, z, zeta, m.area, m.active)
    g64 = ctx.bve_sphere_velocity(x, y, z, zeta, m.area, m.active, precision="fp64")
    g32 = ctx.bve_sphere_velocity(x, y, z, zeta, m.area, m.active, precision="fp32")
    _print("sphere off-sphere", fp64=rel_err_max(g64, ref), fp32=rel_err_max(g32, ref))
    assert rel_err_max(g64, ref) < 1e-10
    assert rel_err_max(g32, ref) < FP32_VEL_TOL


def test_sphere_velocity_radius(ctx, oracle):
    m, zeta = cfg0_inputs(2)
    R = 6.371
    x, y, z = m.x * R, m.y * R, m.z * R
    a = m.area * R * R
    ref = oracle.bve_sphere_velocity(x, y, z, zeta, a, m.active, radius=R)
    got = ctx.bve_sphere_velocity(x, y, z, zeta, a, m.active, radius=R, precision="fp64")
    assert rel_err_max(got, ref) < FP64_TOL


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 5e-7)])
def test_sphere_rank_blocks_concatenate(ctx, prec, tol):
    """MPISetup emulation: P disjoint target blocks concatenate to the full call.
    The source split count follows the block size, so FP32 tile partials (and
    their rounding) differ between the calls; FP64 agrees to rounding."""
    m, zeta = cfg0_inputs(3)
    full = np.array(ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision=prec))
    s, e, _ = load_balance(m.n, 4)
    cat = np.zeros_like(full)
    for r in range(4):
        part = ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, i0=int(s[r]) - 1, i1=int(e[r]),
                                       precision=prec)
        sl = slice(int(s[r]) - 1, int(e[r]))
        cat[:, sl] = np.array(part)[:, sl]
        outside = np.ones(m.n, bool)
        outside[sl] = False
        assert np.all(np.array(part)[:, outside] == 0.0)
    assert rel_err_max(tuple(cat), tuple(full)) < tol


def test_sphere_velocity_deterministic(ctx):
    m, zeta = cfg0_inputs(3)
    a = ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision="fp32")
    b = ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision="fp32")
    for p, q in zip(a, b):
        assert np.array_equal(p, q)


def test_sphere_velocity_solid_body_closed_form(ctx):
    """zeta = 2 w0 z gives u = w0 z_hat x x (BVESolidBody.f90).  The point-vortex
    quadrature of the singular kernel converges at first order in the max norm
    (the error halves per level); a sign, factor or index slip gives O(1)."""
    errs = []
    for level in (4, 5, 6):
        m = icos(level)
        zeta = ic.solid_body_vorticity(m.x, m.y, m.z, 1.0)
        u = np.array(ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision="fp32")).T
        ex = ic.solid_body_velocity(m.x, m.y, m.z, 1.0)
        errs.append(np.abs(u - ex).max() / np.abs(ex).max())
    _print("solid body icos4-6", e4=errs[0], e5=errs[1], e6=errs[2])
    assert errs[2] < 2e-3
    assert errs[0] / errs[1] > 1.7 and errs[1] / errs[2] > 1.7


def test_sphere_velocity_rh4_closed_form_cubed(ctx):
    """RH4 on the cubed sphere (a Y_5^4 + Y_1^0 field): u = -(x cross grad zeta)/(n(n+1))."""
    m = mesh.cubed_sphere_mesh(5)
    zeta = ic.rh4_vorticity(m.x, m.y, m.z)
    u = np.array(ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision="fp32")).T
    ex = ic.rh4_velocity(m.x, m.y, m.z)
    e = np.abs(u - ex).max() / np.abs(ex).max()
    _print("rh4 cubed5", err=e)
    assert e < 2e-2


def test_sphere_velocity_invariants_large(ctx):
    """Size-independent checks at icosahedral L6 (~123k targets): u.x = 0 for
    every target and sum_active zeta A u = 0 (antisymmetric kernel)."""
    m, _ = cfg0_inputs(6)
    rng = np.random.default_rng(3)
    zeta = rng.normal(size=m.n)
    u = np.array(ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision="fp32"))
    X = np.stack([m.x, m.y, m.z])
    umax = np.sqrt((u**2).sum(0)).max()
    assert np.abs((u * X).sum(0)).max() < 1e-6 * umax
    act = m.active == 1
    mom = (zeta[act] * m.area[act] * u[:, act]).sum(1)
    scale = (np.abs(zeta[act] * m.area[act]) * np.sqrt((u[:, act] ** 2).sum(0))).sum()
    assert np.abs(mom).max() < 1e-6 * scale


def test_sphere_velocity_linear_in_zeta(ctx):
    """Scaling zeta by 2 scales every FP32 term exactly: bitwise 2x output."""
    m, zeta = cfg0_inputs(4)
    a = np.array(ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision="fp32"))
    b = np.array(ctx.bve_sphere_velocity(m.x, m.y, m.z, 2.0 * zeta, m.area, m.active, precision="fp32"))
    assert np.array_equal(b, 2.0 * a)


# ----------------------------------------------------------------- plane BS
def _plane_case(level=3, radius=1.0, seed=4):
    m = mesh.quad_rect_mesh(level, half_width=radius)
    rng = np.random.default_rng(seed)
    r2 = m.x**2 + m.y**2
    vort = np.exp(-8.0 * r2) + 0.1 * rng.normal(size=m.n)
    return m, vort


@pytest.mark.parametrize("prec,tol", [("fp64", FP64_TOL), ("fp32", FP32_VEL_TOL)])
def test_plane_velocity(ctx, oracle, prec, tol):
    m, vort = _plane_case(4)
    ref = oracle.plane_velocity(m.x, m.y, vort, m.area, m.active)
    got = ctx.plane_velocity(m.x, m.y, vort, m.area, m.active, precision=prec)
    e = rel_err_max(got, ref)
    _print(f"plane {prec}", err=e)
    assert e < tol


def test_plane_vortex_pair(ctx):
    """Two point vortices: speed Gamma/(2 pi d) perpendicular to the separation."""
    x = np.array([0.0, 0.5])
    y = np.array([0.0, 0.0])
    vort = np.array([2.0, 3.0])
    area = np.array([1.0, 1.0])
    u, v = ctx.plane_velocity(x, y, vort, area, np.ones(2, np.int32), precision="fp64")
    assert abs(u[0]) < 1e-15 and abs(u[1]) < 1e-15
    assert np.isclose(v[0], -3.0 / (2 * np.pi * 0.5), rtol=1e-14)
    assert np.isclose(v[1], 2.0 / (2 * np.pi * 0.5), rtol=1e-14)


# ---------------------------------------------------------------------- PSE
def _pse_case(n_side=64, jitter=1e-3):
    x, y, area, h = mesh.grid_centres(n_side, 1.0, jitter=jitter * 64 / n_side, seed=5)
    f = np.exp(-8.0 * (x**2 + y**2))
    return x, y, f, area, h


@pytest.mark.parametrize("order", [8, 2])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 1e-4)])
def test_pse_laplacian(ctx, oracle, order, prec, tol):
    x, y, f, area, h = _pse_case()
    act = np.ones(x.shape[0], np.int32)
    eps = 2.0 * h
    ref = oracle.pse_plane_laplacian(x, y, f, area, act, eps, order)
    got = ctx.pse_plane_laplacian(x, y, f, area, act, eps, order=order, precision=prec)
    e = rel_err_max(got, ref)
    _print(f"pse{order} {prec}", err=e, rms=rel_err_rms(got, ref))
    assert e < tol


def test_pse_constant_field_is_zero(ctx):
    """Constant f: every term (f_j - f_i) is exactly 0 in every precision
    (the reference's constLap check, SpherePSEConvTest.f90:374-377)."""
    x, y, _, area, h = _pse_case(32)
    f = np.full(x.shape, 3.25)
    act = np.ones(x.shape[0], np.int32)
    for prec in ("fp64", "fp32"):
        for order in (2, 8):
            lap = ctx.pse_plane_laplacian(x, y, f, area, act, 2 * h, order=order, precision=prec)
            assert np.all(lap == 0.0)


# ------------------------------------------------------------ stream funcs
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 2e-6)])
def test_sphere_stream(ctx, oracle, prec, tol):
    m, zeta = cfg0_inputs(3)
    absv = zeta + 2.0 * 2 * np.pi * m.z
    ref = oracle.bve_sphere_stream(m.x, m.y, m.z, zeta, absv, m.area, m.active)
    got = ctx.bve_sphere_stream(m.x, m.y, m.z, zeta, absv, m.area, m.active, precision=prec)
    e0, e1 = rel_err_max(got[0], ref[0]), rel_err_max(got[1], ref[1])
    _print(f"sphere stream {prec}", rel=e0, abs=e1)
    assert e0 < tol and e1 < tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 2e-6)])
def test_plane_stream(ctx, oracle, prec, tol):
    m, vort = _plane_case(4)
    ref = oracle.plane_stream(m.x, m.y, vort, m.area, m.active)
    got = ctx.plane_stream(m.x, m.y, vort, m.area, m.active, precision=prec)
    e = rel_err_max(got, ref)
    _print(f"plane stream {prec}", err=e)
    assert e < tol


# ------------------------------------------------------------------ solvers
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 1e-6)])
def test_bve_solver_matches_oracle_steps(ctx, oracle, prec, tol):
    m, zeta = cfg0_inputs(3)
    omega = 2 * np.pi
    u0 = oracle.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active)
    st = dict(x=m.x, y=m.y, z=m.z, relvort=zeta, u=u0[0], v=u0[1], w=u0[2])
    ref = oracle.bve_timestep(st, m.area, m.active, 1.0, omega, 0.01, nsteps=3)
    from lpmgpu import BVESolver

    s = BVESolver(ctx, m.x, m.y, m.z, zeta, m.area, m.active, radius=1.0, rotation_rate=omega, precision=prec)
    s.timestep(0.01, 3)
    got = s.state()
    ex = rel_err_max((got["x"], got["y"], got["z"]), (ref["x"], ref["y"], ref["z"]))
    ez = rel_err_max(got["relvort"], ref["relvort"])
    eu = rel_err_max((got["u"], got["v"], got["w"]), (ref["u"], ref["v"], ref["w"]))
    _print(f"bve solver {prec}", pos=ex, zeta=ez, vel=eu)
    assert ex < tol and ez < tol and eu < tol
    ke, ens, circ = s.diagnostics()
    rk = oracle.diagnostics(ref["u"], ref["v"], ref["w"], ref["relvort"], m.area, m.active)
    assert np.allclose([ke, ens, circ], rk, rtol=10 * tol, atol=1e-12)
    s.delete()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 1e-6)])
def test_bve_solver_projected_matches_oracle(ctx, oracle, prec, tol):
    """Projected-stage mode (stage inputs and new positions mapped to R x/|x|)
    against the oracle's projected RK4."""
    m, zeta = cfg0_inputs(3)
    omega = 2 * np.pi
    u0 = oracle.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active)
    st = dict(x=m.x, y=m.y, z=m.z, relvort=zeta, u=u0[0], v=u0[1], w=u0[2])
    ref = oracle.bve_timestep(st, m.area, m.active, 1.0, omega, 0.01, nsteps=3, project=True)
    from lpmgpu import BVESolver

    s = BVESolver(ctx, m.x, m.y, m.z, zeta, m.area, m.active, rotation_rate=omega, precision=prec, project=True)
    s.timestep(0.01, 3)
    got = s.state()
    r = np.sqrt(got["x"] ** 2 + got["y"] ** 2 + got["z"] ** 2)
    assert np.abs(r - 1.0).max() < 1e-14
    ex = rel_err_max((got["x"], got["y"], got["z"]), (ref["x"], ref["y"], ref["z"]))
    eu = rel_err_max((got["u"], got["v"], got["w"]), (ref["u"], ref["v"], ref["w"]))
    _print(f"bve solver projected {prec}", pos=ex, vel=eu)
    assert ex < tol and eu < tol
    s.delete()


def test_bve_solver_fine_mesh_conserves(ctx):
    """Projected RK4 at icosahedral L6 with the stage displacement larger than
    the particle spacing: circulation, enstrophy and KE stay put over 5 steps."""
    m, zeta = cfg0_inputs(6)
    zeta = zeta + ic.rh4_vorticity(m.x, m.y, m.z)
    from lpmgpu import BVESolver

    s = BVESolver(ctx, m.x, m.y, m.z, zeta, m.area, m.active, rotation_rate=2 * np.pi, precision="fp32",
                  project=True)
    d0 = np.array(s.diagnostics())
    s.timestep(0.01, 5)
    d1 = np.array(s.diagnostics())
    _print("fine mesh diag", ke0=d0[0], ke1=d1[0], ens0=d0[1], ens1=d1[1], circ0=d0[2], circ1=d1[2])
    assert np.all(np.isfinite(d1))
    assert abs(d1[0] - d0[0]) < 1e-2 * abs(d0[0])
    assert abs(d1[1] - d0[1]) < 5e-2 * abs(d0[1])
    s.delete()


def test_bve_solver_graph_replay_matches_eager(ctx):
    m, zeta = cfg0_inputs(3)
    from lpmgpu import BVESolver

    a = BVESolver(ctx, m.x, m.y, m.z, zeta, m.area, m.active, rotation_rate=2 * np.pi, precision="fp32")
    b = BVESolver(ctx, m.x, m.y, m.z, zeta, m.area, m.active, rotation_rate=2 * np.pi, precision="fp32",
                  use_graph=True)
    for _ in range(4):  # eager work and synchronizations between replays
        a.timestep(0.01)
        b.timestep(0.01)
        sa, sb = a.state(), b.state()
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), k
    a.delete()
    b.delete()


def test_bve_solver_graph_survives_diagnostics_and_stream_calls():
    """A captured step graph with stream functions in every step, interleaved
    with diagnostics and on-demand stream-function calls on the same context
    (which grow the context's reduction buffer and may regrow the solver's
    workspaces), stays bitwise equal to an eager solver doing the same calls.
    A fresh context, so its buffers start small."""
    import lpmgpu
    from lpmgpu import BVESolver

    m, zeta = cfg0_inputs(4)
    absv = zeta + 2 * (2 * np.pi) * m.z
    with lpmgpu.Context(0) as c:
        kw = dict(absvort=absv, rotation_rate=2 * np.pi, precision="fp32", stream_each_step=True, project=True)
        a = BVESolver(c, m.x, m.y, m.z, zeta, m.area, m.active, **kw)
        b = BVESolver(c, m.x, m.y, m.z, zeta, m.area, m.active, use_graph=True, **kw)
        for k, act in enumerate(["diag", "stream", "diag", "none", "stream"]):
            a.timestep(0.01)
            b.timestep(0.01)
            if act == "diag":
                assert a.diagnostics() == b.diagnostics()
            elif act == "stream":
                pa, pb = a.stream(), b.stream()
                assert all(np.array_equal(u, v) for u, v in zip(pa, pb))
            sa, sb = a.state(), b.state()
            for key in sa:
                assert np.array_equal(sa[key], sb[key]), (k, key)
        a.delete()
        b.delete()


def test_plane_solver_graph_survives_stream_calls():
    import lpmgpu
    from lpmgpu import PlaneSolver

    m, vort = _plane_case(3)
    with lpmgpu.Context(0) as c:
        a = PlaneSolver(c, m.x, m.y, vort, m.area, m.active, precision="fp32", stream_each_step=True)
        b = PlaneSolver(c, m.x, m.y, vort, m.area, m.active, precision="fp32", stream_each_step=True,
                        use_graph=True)
        for k in range(4):
            a.timestep(0.005)
            b.timestep(0.005)
            assert a.diagnostics() == b.diagnostics()
            assert np.array_equal(a.stream(), b.stream())
            sa, sb = a.state(), b.state()
            for key in sa:
                assert np.array_equal(sa[key], sb[key]), (k, key)
        a.delete()
        b.delete()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 1e-6)])
def test_plane_solver_matches_oracle_steps(ctx, oracle, prec, tol):
    m, vort = _plane_case(3)
    u0 = oracle.plane_velocity(m.x, m.y, vort, m.area, m.active)
    st = dict(x=m.x, y=m.y, u=u0[0], v=u0[1])
    ref = oracle.plane_timestep(st, vort, m.area, m.active, 0.005, nsteps=3)
    from lpmgpu import PlaneSolver

    s = PlaneSolver(ctx, m.x, m.y, vort, m.area, m.active, precision=prec)
    s.timestep(0.005, 3)
    got = s.state()
    ex = rel_err_max((got["x"], got["y"]), (ref["x"], ref["y"]))
    eu = rel_err_max((got["u"], got["v"]), (ref["u"], ref["v"]))
    _print(f"plane solver {prec}", pos=ex, vel=eu)
    assert ex < tol and eu < tol
    s.delete()


def test_diagnostics_match_oracle(ctx, oracle):
    m, zeta = cfg0_inputs(3)
    rng = np.random.default_rng(6)
    u, v, w = rng.normal(size=(3, m.n))
    got = ctx.diagnostics(u, v, w, zeta, m.area, m.active)
    ref = oracle.diagnostics(u, v, w, zeta, m.area, m.active)
    assert np.allclose(got, ref, rtol=1e-13, atol=0)


# --------------------------------------------------------------- error paths
def test_invalid_arguments_report_errors(ctx):
    m, zeta = cfg0_inputs(1)
    with pytest.raises(LpmError) as e:
        ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, i0=5, i1=2)
    assert e.value.code == 1 and "i0" in str(e.value)
    with pytest.raises(LpmError) as e:
        ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision=7)
    assert e.value.code == 1
    with pytest.raises(LpmError):
        ctx.pse_plane_laplacian(m.x, m.y, zeta, m.area, m.active, eps=0.1, order=4)
    # the context stays usable after errors
    u = ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision="fp64")
    assert np.isfinite(np.array(u)).all()


def test_empty_blocks_and_no_particles(ctx):
    """An empty target block [i0, i0) leaves every output untouched (zeros from
    the wrapper) and n = 0 returns empty arrays, for each drop-in sum, with the
    context usable afterwards."""
    m, zeta = cfg0_inputs(1)
    absv = zeta + 2.0 * 2 * np.pi * m.z
    e = np.zeros(0)
    none = np.zeros(0, np.int32)
    for prec in ("fp64", "fp32"):
        calls = (
            lambda **b: ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision=prec, **b),
            lambda **b: ctx.plane_velocity(m.x, m.y, zeta, m.area, m.active, precision=prec, **b),
            lambda **b: ctx.pse_plane_laplacian(m.x, m.y, zeta, m.area, m.active, 0.1, precision=prec, **b),
            lambda **b: ctx.bve_sphere_stream(m.x, m.y, m.z, zeta, absv, m.area, m.active, precision=prec, **b),
            lambda **b: ctx.plane_stream(m.x, m.y, zeta, m.area, m.active, precision=prec, **b),
        )
        for k, call in enumerate(calls):
            out = np.array(call(i0=7, i1=7))
            assert out.shape[-1] == m.n and not out.any(), (prec, k)
        empty = (
            ctx.bve_sphere_velocity(e, e, e, e, e, none, precision=prec),
            ctx.plane_velocity(e, e, e, e, none, precision=prec),
            ctx.pse_plane_laplacian(e, e, e, e, none, 0.1, precision=prec),
            ctx.bve_sphere_stream(e, e, e, e, e, e, none, precision=prec),
            ctx.plane_stream(e, e, e, e, none, precision=prec),
        )
        for k, out in enumerate(empty):
            assert np.array(out).size == 0, (prec, k)
    u = ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active, precision="fp64")
    assert np.isfinite(np.array(u)).all() and np.array(u).any()


def test_no_active_sources_gives_zero(ctx):
    m, zeta = cfg0_inputs(1)
    u = ctx.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, np.zeros(m.n, np.int32), precision="fp32")
    assert all(np.all(c == 0.0) for c in u)
