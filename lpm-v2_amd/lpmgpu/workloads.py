"""The five measured workloads as concrete, seeded inputs.

  cfg0  icosahedral L4 single Gaussian vortex, FP64, dt 0.01 x 10 steps
  cfg1  cubed sphere L7 Rossby-Haurwitz 4, FP32-compensated pairs, dt 0.005 x 100
  cfg2  512^2 jittered grid, PSE Laplacian of exp(-8 r^2), eps = 2h
  cfg3  64 icosahedral problems, levels U{3..7}, 1-4 Gaussian vortices each
  cfg4  icosahedral L8, Gaussian vortex + RH4 background (learned far field)
"""
from __future__ import annotations

import numpy as np

from . import ic, mesh

BASE_SEED = 20261202
OMEGA = 2.0 * np.pi


def cfg0():
    m = mesh.icosahedral_mesh(4)
    zeta = ic.gaussian_vortex(m.x, m.y, m.z, ic.lat_lon_point(0.3, 0.0), 4.0 * np.pi, 16.0)
    return dict(mesh=m, relvort=zeta, absvort=zeta + 2 * OMEGA * m.z, dt=0.01, steps=10, omega=OMEGA)


def cfg1(level=7):
    m = mesh.cubed_sphere_mesh(level)
    zeta = ic.rh4_vorticity(m.x, m.y, m.z)
    return dict(mesh=m, relvort=zeta, absvort=zeta + 2 * OMEGA * m.z, dt=0.005, steps=100, omega=OMEGA)


def cfg2(n_side=512, seed=BASE_SEED, jitter=1e-3):
    x, y, area, h = mesh.grid_centres(n_side, 1.0, jitter=jitter, seed=seed)
    r2 = x * x + y * y
    f = np.exp(-8.0 * r2)
    exact = (256.0 * r2 - 32.0) * np.exp(-8.0 * r2)
    return dict(x=x, y=y, area=area, f=f, exact=exact, h=h, eps=2.0 * h)


def cfg3(n_inst=64, levels=(3, 7), seed=BASE_SEED, max_level=None):
    """64 problems; level drawn U{3..7}; 1-4 Gaussian vortices with strength
    U(-4 pi, 4 pi), width U(8, 32) in exp(-width |x - c|^2), centres uniform on
    the sphere; Omega = 2 pi.  max_level caps the draw (for quick tests)."""
    rng = np.random.default_rng(seed)
    lev = rng.integers(levels[0], levels[1] + 1, size=n_inst)
    if max_level is not None:
        lev = np.minimum(lev, max_level)
    cache = {}
    probs = []
    for k in range(n_inst):
        L = int(lev[k])
        if L not in cache:
            cache[L] = mesh.icosahedral_mesh(L)
        m = cache[L]
        nv = int(rng.integers(1, 5))
        zeta = ic.random_gaussian_vortices(rng, nv, m.x, m.y, m.z)
        probs.append(dict(x=m.x, y=m.y, z=m.z, relvort=zeta, area=m.area, active=m.active, radius=1.0,
                          rotation_rate=OMEGA, level=L, n_vortex=nv))
    return dict(problems=probs, dt=0.01, steps=20)


def cfg4(level=8, cell_level=3):
    m = mesh.icosahedral_mesh(level, cell_level=cell_level)
    zeta = ic.gaussian_vortex(m.x, m.y, m.z, ic.lat_lon_point(0.3, 0.0), 4.0 * np.pi, 16.0) + ic.rh4_vorticity(
        m.x, m.y, m.z)
    return dict(mesh=m, relvort=zeta, absvort=zeta + 2 * OMEGA * m.z, dt=0.01, steps=100, omega=OMEGA)
