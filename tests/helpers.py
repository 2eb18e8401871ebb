"""Shared inputs and error metrics for the tests."""
import functools

import numpy as np

from lpmgpu import ic, mesh


@functools.lru_cache(maxsize=None)
def icos(level, cell_level=-1):
    return mesh.icosahedral_mesh(level, cell_level=cell_level)


@functools.lru_cache(maxsize=None)
def cubed(level):
    return mesh.cubed_sphere_mesh(level)


def cfg0_inputs(level=4):
    """Gaussian vortex 4 pi exp(-16 |x - x0|^2), x0 at lat 0.3, lon 0 (relative vorticity)."""
    m = icos(level)
    x0 = ic.lat_lon_point(0.3, 0.0)
    zeta = ic.gaussian_vortex(m.x, m.y, m.z, x0, 4.0 * np.pi, 16.0)
    return m, zeta


def rh4_inputs(m):
    return ic.rh4_vorticity(m.x, m.y, m.z)


def rel_err_max(test, ref):
    """Normwise max relative error max_i |d_i| / max_i |ref_i|; vectors are given
    as a tuple of component arrays."""
    t = np.atleast_2d(np.asarray(test, dtype=np.float64))
    r = np.atleast_2d(np.asarray(ref, dtype=np.float64))
    d = np.sqrt(np.sum((t - r) ** 2, axis=0))
    m = np.sqrt(np.sum(r**2, axis=0))
    return float(d.max() / m.max())


def taylor_cell_sum(w, Y, c, x, order, npts=64):
    """Order-`order` Taylor polynomial, at t = 1, of h(t) = sum_j w_j y_j(t) / (1 - x.y_j(t))
    with y_j(t) = c + t (Y_j - c) (unit sphere, |x| = 1): the cell's multipole
    expansion about c.  The coefficients come from Cauchy's integral on the circle
    |t| = rho inside the radius of convergence, so no moment or derivative
    formula is involved."""
    d = np.asarray(Y, np.float64) - c
    xd = d @ x
    D = 1.0 - c @ x
    big = np.abs(xd).max()
    rho = 1.0 if big == 0 else min(1.0, 0.5 * D / big)
    th = 2 * np.pi * np.arange(npts) / npts
    t = rho * np.exp(1j * th)
    H = np.einsum("j,mjk->mk", w, (c[None, None, :] + t[:, None, None] * d[None, :, :])
                  / (D - t[:, None] * xd[None, :])[:, :, None])
    out = np.zeros(3)
    for n in range(order + 1):
        out += (np.exp(-1j * n * th) @ H).real / npts / rho ** n
    return out


def rel_err_rms(test, ref):
    t = np.atleast_2d(np.asarray(test, dtype=np.float64))
    r = np.atleast_2d(np.asarray(ref, dtype=np.float64))
    return float(np.sqrt(np.sum((t - r) ** 2) / np.sum(r**2)))
