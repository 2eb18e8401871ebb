"""The oracle and the mesh generator against golden vectors written by
lpm-v2's own Fortran (tests/golden/ref_*.npz; regenerate with
tests/golden/make_golden.py, which builds the reference's modules with
oracle/ref.mk).

Cases: icosahedral L2 and L3 with the Gaussian vortex, cubed sphere L3 with
the Rossby-Haurwitz-4 wave, quadRect L3 with two planar Gaussians.  Each
holds the reference's particles (positions, areas, active flags, in its
creation order), its SetVelocityOnMesh and stream-function outputs, one
Timestep of its RK4 solver, and (planar) its PSE Laplacian and gradient.
"""
import os

import numpy as np
import pytest

from lpmgpu import mesh

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")

# The oracle restates the reference's loops in the same FP64 operation order,
# but the C compiler and amdflang may contract or reorder differently:
# agreement is to rounding.
TOL = 1e-13


def load(case):
    with np.load(os.path.join(GOLDEN, f"ref_{case}.npz")) as g:
        return {k: g[k] for k in g.files}


def rel(test, ref):
    t = np.atleast_2d(np.asarray(test, dtype=np.float64))
    r = np.atleast_2d(np.asarray(ref, dtype=np.float64))
    return float(np.sqrt(((t - r) ** 2).sum(0)).max() / np.sqrt((r**2).sum(0)).max())


SPHERE_MESHES = {"icos2": lambda: mesh.icosahedral_mesh(2), "icos3": lambda: mesh.icosahedral_mesh(3),
                 "cubed3": lambda: mesh.cubed_sphere_mesh(3)}


@pytest.mark.parametrize("case", sorted(SPHERE_MESHES))
def test_sphere_mesh_matches_reference(case):
    g = load(case)
    m = SPHERE_MESHES[case]()
    assert m.n == int(g["n"])
    assert np.abs(np.stack([m.x, m.y, m.z]) - np.stack([g["x"], g["y"], g["z"]])).max() < 1e-15
    assert np.abs(m.area - g["area"]).max() < 1e-15
    assert np.array_equal(m.active.astype(np.float64), g["active"])


def test_plane_mesh_matches_reference():
    g = load("quad3")
    m = mesh.quad_rect_mesh(3)
    assert m.n == int(g["n"])
    assert np.abs(np.stack([m.x, m.y]) - np.stack([g["x"], g["y"]])).max() < 1e-15
    assert np.abs(m.area - g["area"]).max() < 1e-15
    assert np.array_equal(m.active.astype(np.float64), g["active"])


@pytest.mark.parametrize("case", sorted(SPHERE_MESHES))
def test_oracle_sphere_sums_match_reference(oracle, case):
    g = load(case)
    act = g["active"].astype(np.int32)
    u = oracle.bve_sphere_velocity(g["x"], g["y"], g["z"], g["relvort"], g["area"], act)
    assert rel(u, (g["u"], g["v"], g["w"])) < TOL
    pr, pa = oracle.bve_sphere_stream(g["x"], g["y"], g["z"], g["relvort"], g["absvort"], g["area"], act)
    assert rel(pr, g["psi_rel"]) < TOL
    assert rel(pa, g["psi_abs"]) < TOL
    # the reference's initial absolute vorticity is zeta + 2 Omega z / R (SphereBVE.f90:345)
    assert np.abs(g["absvort"] - (g["relvort"] + 2 * (2 * np.pi) * g["z"])).max() < 1e-13


def test_oracle_sphere_step_matches_reference(oracle):
    g = load("icos2")
    act = g["active"].astype(np.int32)
    st = dict(x=g["x"], y=g["y"], z=g["z"], relvort=g["relvort"], u=g["u"], v=g["v"], w=g["w"])
    o = oracle.bve_timestep(st, g["area"], act, 1.0, 2 * np.pi, 0.01)
    assert rel((o["x"], o["y"], o["z"]), (g["step1_x"], g["step1_y"], g["step1_z"])) < TOL
    assert rel(o["relvort"], g["step1_relvort"]) < TOL
    assert rel((o["u"], o["v"], o["w"]), (g["step1_u"], g["step1_v"], g["step1_w"])) < TOL
    # the reference ends its step with the stream functions (SphereBVESolver.f90:352),
    # psi_abs from the set-up absolute vorticity, which the step leaves unchanged
    pr, pa = oracle.bve_sphere_stream(o["x"], o["y"], o["z"], o["relvort"], g["absvort"], g["area"], act)
    assert rel(pr, g["step1_psi_rel"]) < TOL
    assert rel(pa, g["step1_psi_abs"]) < TOL


def test_oracle_plane_sums_and_step_match_reference(oracle):
    g = load("quad3")
    act = g["active"].astype(np.int32)
    assert rel(oracle.plane_velocity(g["x"], g["y"], g["vort"], g["area"], act), (g["u"], g["v"])) < TOL
    assert rel(oracle.plane_stream(g["x"], g["y"], g["vort"], g["area"], act), g["psi"]) < TOL
    st = dict(x=g["x"], y=g["y"], u=g["u"], v=g["v"])
    o = oracle.plane_timestep(st, g["vort"], g["area"], act, 0.005)
    assert rel((o["x"], o["y"]), (g["step1_x"], g["step1_y"])) < TOL
    assert rel((o["u"], o["v"]), (g["step1_u"], g["step1_v"])) < TOL
    assert rel(oracle.plane_stream(o["x"], o["y"], g["vort"], g["area"], act), g["step1_psi"]) < TOL


def test_oracle_pse_matches_reference(oracle):
    """PSE Laplacian and gradient with the reference's own eps = MaxEdgeLength^0.75
    (PSEDirectSum.f90:93-113): h = 0.125 on quadRect L3 over [-1, 1]^2."""
    g = load("quad3")
    act = g["active"].astype(np.int32)
    eps = float(g["pse_eps"][0])
    assert abs(eps - oracle.pse_eps(0.125, 0.75)) < 1e-15
    lap = oracle.pse_plane_laplacian(g["x"], g["y"], g["pse_f"], g["area"], act, eps, 8)
    assert rel(lap, g["pse_lap"]) < TOL
    grad = oracle.pse_plane_gradient(g["x"], g["y"], g["pse_f"], g["area"], act, eps)
    assert rel(grad, (g["pse_gx"], g["pse_gy"])) < TOL


def test_golden_catches_a_planted_mistake(oracle):
    """A sign slip in one source's weight (a plausible restatement error) moves
    the velocity far outside the tolerance the tests above hold."""
    g = load("icos3")
    act = g["active"].astype(np.int32)
    q = g["relvort"].copy()
    j = int(np.argmax(np.abs(q * g["area"] * act)))  # the strongest source
    q[j] = -q[j]
    u = oracle.bve_sphere_velocity(g["x"], g["y"], g["z"], q, g["area"], act)
    assert rel(u, (g["u"], g["v"], g["w"])) > 1e3 * TOL


@pytest.mark.skipif(not (os.path.exists(os.path.join(ROOT, "oracle", "_ref", "ref_cpu_bench"))
                         and os.path.exists("/opt/conda/bin/mpiexec")),
                    reason="the reference's CPU-baseline driver is built only where its sources are "
                           "(make -f oracle/ref.mk)")
def test_reference_cpu_baseline_runs_the_reference_loop():
    """bench.py's CPU baseline runs lpm-v2's own SetVelocityOnMesh on MPI ranks
    over a block of targets; the block it computes equals the oracle's."""
    import sys

    sys.path.insert(0, ROOT)
    import bench
    from lpmgpu import ic

    m = mesh.icosahedral_mesh(4)
    z = ic.gaussian_vortex(m.x, m.y, m.z, ic.lat_lon_point(0.3, 0.0), 4 * np.pi, 16.0) + \
        ic.rh4_vorticity(m.x, m.y, m.z)
    out = bench.reference_cpu_baseline(m, z, 0.5, 2)
    assert out is not None and out["kind"] == "reference" and out["cores"] == 2
    assert out["value"] > 0
    assert out["max_rel_diff_vs_oracle"] < 1e-12
