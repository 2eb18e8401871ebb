"""CPU checks of the numpy references the far-field GPU tests rely on."""
import numpy as np

from helpers import taylor_cell_sum
from lpmgpu import ic, mesh
from lpmgpu.farfield import FarFieldBVE


def _unit(v):
    return v / np.linalg.norm(v)


def test_taylor_single_source_is_geometric_series():
    """One source y = c + d: y / (D - x.d) = (c + d) phi sum_n (phi x.d)^n, so the
    order-P polynomial is known in closed form."""
    rng = np.random.default_rng(1)
    for _ in range(20):
        c = 0.99 * _unit(rng.normal(size=3))
        y = _unit(c + 0.05 * rng.normal(size=3))
        x = _unit(c + 0.4 * rng.normal(size=3))
        d = y - c
        D = 1 - x @ c
        phi = 1 / D
        q = phi * (x @ d)
        for order in range(0, 6):
            want = c * phi * sum(q ** n for n in range(order + 1)) + d * phi * sum(q ** n for n in range(order))
            got = taylor_cell_sum(np.array([1.0]), y[None, :], c, x, order)
            assert np.allclose(got, want, rtol=1e-11, atol=1e-12 * np.abs(want).max())


def test_taylor_cell_converges_to_the_exact_sum():
    """For a level-3 cell of a level-5 mesh and a target 0.5 rad away the error of
    the order-P expansion falls by a roughly constant factor per order."""
    m = mesh.icosahedral_mesh(5, cell_level=3)
    cells = FarFieldBVE.particle_cells(m)
    zeta = ic.gaussian_vortex(m.x, m.y, m.z, ic.lat_lon_point(0.3, 0.0), 4 * np.pi, 16.0)
    P = m.xyz()
    idx = np.flatnonzero((m.active == 1) & (cells == 7))
    w = (zeta * m.area)[idx]
    c = (P[idx] * m.area[idx, None]).sum(0) / m.area[idx].sum()
    t = _unit(np.cross(c, [0.0, 0.0, 1.0]))
    x = np.cos(0.5) * _unit(c) + np.sin(0.5) * t
    exact = (w[:, None] * P[idx] / (1 - P[idx] @ x)[:, None]).sum(0)
    err = [np.linalg.norm(taylor_cell_sum(w, P[idx], c, x, k) - exact) / np.linalg.norm(exact) for k in range(9)]
    # the ratio is about l / (2 |x - c|) ~ 1/4 here
    assert all(0.1 < err[k + 1] / err[k] < 0.5 for k in range(8)), err
    assert err[8] < 2e-6
