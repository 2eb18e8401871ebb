"""Remesh re-upload: lpm_bve_solver_set_particles, lpm_plane_solver_set_particles
and lpm_bve_ff_set_particles replace a solver's particle set (the output of
LagrangianRemeshBVEWithVorticityFunction, SsrfpackRemesh.f90:388-, every
remeshInterval steps in BVESingleGaussianVortex.f90:218-246) and recompute the
velocity as SetVelocityOnMesh does after a remesh (SsrfpackRemesh.f90:552).

A reloaded solver must behave exactly like a solver created on the new set:
same state, bit for bit, after the same steps, for larger and smaller sets,
with and without a captured step graph."""
import time

import numpy as np
import pytest

import lpmgpu
from helpers import icos, rel_err_max
from lpmgpu import ic, mesh
from lpmgpu.farfield import EXPANSION, PLAIN, FarFieldBVE

pytestmark = pytest.mark.gpu


def _vort(m):
    """The vorticity function the remesh interpolates: a Gaussian vortex + RH4."""
    return ic.gaussian_vortex(m.x, m.y, m.z, ic.lat_lon_point(0.3, 0.0), 4 * np.pi, 16.0) + \
        ic.rh4_vorticity(m.x, m.y, m.z)


def _args(m):
    return (m.x, m.y, m.z, _vort(m), m.area, m.active)


def _same_state(a, b):
    return all(np.array_equal(a[k], b[k]) for k in a)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("use_graph", [False, True])
def test_bve_reload_matches_fresh_solver(ctx, precision, use_graph):
    """L4 -> (2 steps) -> reload L5 -> 2 steps, then reload L3 -> 2 steps; each
    leg equals a solver created on that set."""
    kw = dict(radius=1.0, rotation_rate=2 * np.pi, precision=precision, use_graph=use_graph, project=True)
    s = lpmgpu.BVESolver(ctx, *_args(icos(4)), **kw)
    s.timestep(0.01, 2)
    for lev in (5, 3):
        m = icos(lev)
        s.set_particles(*_args(m))
        assert s.n == m.n
        fresh = lpmgpu.BVESolver(ctx, *_args(m), **kw)
        assert _same_state(s.state(), fresh.state()), f"velocity after reload to L{lev}"
        s.timestep(0.01, 2)
        fresh.timestep(0.01, 2)
        assert _same_state(s.state(), fresh.state()), f"state after 2 steps on L{lev}"
        fresh.delete()
    s.delete()


def test_bve_reload_velocity_is_the_direct_sum(ctx):
    """Without u, v, w the reload computes the velocity on the device: it is
    the sphere Biot-Savart sum of the new set."""
    s = lpmgpu.BVESolver(ctx, *_args(icos(3)), precision="fp64")
    m = icos(4)
    s.set_particles(*_args(m))
    st = s.state()
    ref = ctx.bve_sphere_velocity(m.x, m.y, m.z, _vort(m), m.area, m.active, precision="fp64")
    assert rel_err_max((st["u"], st["v"], st["w"]), ref) < 1e-13
    # given velocities are taken as they are
    u = np.full(m.n, 0.25)
    s.set_particles(*_args(m), u=u, v=-u, w=2 * u)
    st = s.state()
    assert np.array_equal(st["u"], u) and np.array_equal(st["w"], 2 * u)
    s.delete()


def test_bve_reload_l8(ctx):
    """Reload at icosahedral L8 (1,966,082 particles): with the velocity given,
    the state holds exactly what was passed; without it, the velocity the
    solver recomputes on the device equals the standalone FP32 sum of the same
    particles within the FP32 velocity bar.  Reload times are reported, not
    asserted."""
    m = icos(8)
    args = _args(m)
    rng = np.random.default_rng(8)
    ug, vg, wg = (rng.standard_normal(m.n) for _ in range(3))
    z = np.zeros(m.n)
    s = lpmgpu.BVESolver(ctx, *args, u=z, v=z, w=z, precision="fp32", project=True)
    t0 = time.perf_counter()
    s.set_particles(*args, u=ug, v=vg, w=wg)
    t_copy = time.perf_counter() - t0
    st = s.state()
    assert np.array_equal(st["u"], ug) and np.array_equal(st["v"], vg) and np.array_equal(st["w"], wg)
    assert np.abs(st["x"] - args[0]).max() < 1e-15  # projected onto the sphere at load: at most an ulp
    assert np.array_equal(st["relvort"], args[3])
    t0 = time.perf_counter()
    s.set_particles(*args)
    t_vel = time.perf_counter() - t0
    st = s.state()
    print(f"L8 reload: {1e3 * t_copy:.1f} ms with the velocity given, {1e3 * t_vel:.1f} ms with the velocity sum")
    s.delete()
    ref = np.stack(ctx.bve_sphere_velocity(m.x, m.y, m.z, args[3], m.area, m.active, precision="fp32"))
    got = np.stack([st["u"], st["v"], st["w"]])
    e = np.sqrt(((got - ref) ** 2).sum(0)).max() / np.sqrt((ref**2).sum(0)).max()
    print(f"L8 reload velocity vs standalone sum: {e:.2e}")
    assert e < 1e-6


@pytest.mark.parametrize("use_graph", [False, True])
def test_plane_reload_matches_fresh_solver(ctx, use_graph):
    def plane(n_side, seed):
        x, y, area, h = mesh.grid_centres(n_side, 1.0, jitter=0.2 * 2.0 / n_side, seed=seed)
        vort = 4.0 * np.exp(-6 * ((x - 0.3) ** 2 + y ** 2)) - 3.0 * np.exp(-6 * ((x + 0.3) ** 2 + y ** 2))
        act = np.ones(x.size, np.int32)
        act[::11] = 0
        return x, y, vort, area, act

    s = lpmgpu.PlaneSolver(ctx, *plane(32, 1), precision="fp32", use_graph=use_graph)
    s.timestep(0.01, 2)
    for side in (48, 24):
        a = plane(side, side)
        s.set_particles(*a)
        fresh = lpmgpu.PlaneSolver(ctx, *a, precision="fp32", use_graph=use_graph)
        assert _same_state(s.state(), fresh.state())
        s.timestep(0.01, 2)
        fresh.timestep(0.01, 2)
        assert _same_state(s.state(), fresh.state())
        fresh.delete()
    s.delete()


def test_reload_errors_leave_the_solver_usable(ctx):
    m = icos(3)
    s = lpmgpu.BVESolver(ctx, *_args(m), precision="fp64")
    with pytest.raises(lpmgpu.LpmError):
        s.set_particles(m.x[:0], m.y[:0], m.z[:0], m.x[:0], m.area[:0], m.active[:0])
    with pytest.raises(lpmgpu.LpmError):
        s.set_particles(*_args(m), u=np.zeros(m.n))  # u without v, w
    s.timestep(0.01)  # still usable: argument errors change nothing
    s.delete()


@pytest.mark.parametrize("parent", [0.0, 0.9])
def test_far_field_reload_keeps_cells_and_network(ctx, parent):
    """The far field keeps its cells, its coarse level and its network across a
    reload to a finer set; its plain velocity is the direct sum of the new set
    and the expansion stays within its usual error."""
    m4 = mesh.icosahedral_mesh(4, cell_level=2)
    f = FarFieldBVE(ctx, m4, _vort(m4), near_radius=0.3, band_radius=0.6, parent_radius=parent)
    p0 = f.params()
    m5 = mesh.icosahedral_mesh(5, cell_level=2)
    zeta = _vort(m5)
    f.set_particles(m5.x, m5.y, m5.z, zeta, m5.area, m5.active)
    assert f.n == m5.n and f.state()["x"].shape == (m5.n,)
    assert np.array_equal(f.params(), p0)
    plain = f.velocity(PLAIN)
    ref = ctx.bve_sphere_velocity(m5.x, m5.y, m5.z, zeta, m5.area, m5.active, precision="fp64")
    e_plain = rel_err_max(plain, ref)
    e_exp = rel_err_max(f.velocity(EXPANSION), plain)
    fresh = FarFieldBVE(ctx, m5, zeta, near_radius=0.3, band_radius=0.6, parent_radius=parent)
    e_fresh = rel_err_max(fresh.velocity(EXPANSION), fresh.velocity(PLAIN))
    if parent > 0:  # the coarse level is in use after the reload: fewer records than non-near cells
        for g in (f, fresh):
            near, _ = g.cells()
            assert (g.far_counts() + near).mean() < g.ncell - 1
    print("ff reload: plain vs fp64", e_plain, "expansion", e_exp, "fresh object", e_fresh)
    assert e_plain < 1e-6
    assert e_exp < 3 * e_fresh + 1e-7
    f.steps(0.01, 2, EXPANSION, judge_every=0)
    fresh.delete()
    f.delete()
