"""Initial vorticity fields and closed-form velocities on the unit-radius sphere.

* Gaussian vortex: `GaussianVortexVorticity` (BVESingleGaussianVortex.f90:334-340)
  has the form strength * exp(-beta^2 (R^2 - x.c)) - GAUSS_CONST; the
  workloads here state it as strength * exp(-width |x - c|^2), which is the
  same family since R^2 - x.c = |x - c|^2 / 2 on the sphere.  The optional
  mean removal restates `SetGaussConst` (BVESingleGaussianVortex.f90:350-357).
* Rossby-Haurwitz 4: relative vorticity 2 w z - 30 K z (x^4 - 6 x^2 y^2 + y^4)
  (= 2 w sin(phi) - 30 K sin(phi) cos^4(phi) cos(4 lambda)); the reference's
  `RossbyHaurwitz54Vorticity` (RossbyHaurwitz54.f90:413-428) uses the
  opposite sign on the wave term.
* Solid body: zeta = 2 w0 z / R with velocity w0 z_hat x x (BVESolidBody.f90).

Absolute vorticity is relative vorticity plus 2 Omega z / R (SphereBVE.f90:347);
the Biot-Savart sum consumes relative vorticity (SphereBVESolver.f90:257-259).
"""
from __future__ import annotations

import numpy as np


def lat_lon_point(lat: float, lon: float) -> np.ndarray:
    return np.array([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)])


def gaussian_vortex(x, y, z, centre, strength=4.0 * np.pi, width=16.0):
    d2 = (x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2
    return strength * np.exp(-width * d2)


def remove_mean(zeta, area, active, radius=1.0):
    """zeta - sum(zeta A)/(4 pi R) over active particles (SetGaussConst)."""
    c = np.sum(zeta * area * (active != 0)) / (4.0 * np.pi * radius)
    return zeta - c


def rh4_vorticity(x, y, z, omega=1.0, K=1.0):
    return 2.0 * omega * z - 30.0 * K * z * (x**4 - 6.0 * x**2 * y**2 + y**4)


def rh4_velocity(x, y, z, omega=1.0, K=1.0):
    """Closed-form Biot-Savart velocity of `rh4_vorticity` on the unit sphere.

    For a zero-mean degree-n spherical harmonic zeta = P|_S (P a homogeneous
    harmonic polynomial), the stream function is -P / (n (n + 1)) and the
    velocity is x cross grad_s psi = -(x cross grad P) / (n (n + 1)).
    """
    # n = 1 part, P1 = 2 w z: u = w z_hat x x
    u1 = np.stack([-omega * y, omega * x, np.zeros_like(x)], axis=1)
    # n = 5 part, P5 = -30 K z (x^4 - 6 x^2 y^2 + y^4)
    q = x**4 - 6.0 * x**2 * y**2 + y**4
    gx = -30.0 * K * z * (4.0 * x**3 - 12.0 * x * y**2)
    gy = -30.0 * K * z * (-12.0 * x**2 * y + 4.0 * y**3)
    gz = -30.0 * K * q
    g = np.stack([gx, gy, gz], axis=1)
    X = np.stack([x, y, z], axis=1)
    u5 = -np.cross(X, g) / 30.0
    return u1 + u5


def solid_body_vorticity(x, y, z, w0=1.0, radius=1.0):
    return 2.0 * w0 * z / radius


def solid_body_velocity(x, y, z, w0=1.0):
    return np.stack([-w0 * y, w0 * x, np.zeros_like(x)], axis=1)


def random_gaussian_vortices(rng, n_vortex, x, y, z):
    """Sum of n Gaussian vortices, strength U(-4 pi, 4 pi), width U(8, 32),
    centres uniform on the sphere."""
    zeta = np.zeros_like(x)
    for _ in range(n_vortex):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        s = rng.uniform(-4.0 * np.pi, 4.0 * np.pi)
        w = rng.uniform(8.0, 32.0)
        zeta += gaussian_vortex(x, y, z, v, s, w)
    return zeta
