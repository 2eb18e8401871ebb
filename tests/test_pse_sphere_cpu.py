"""CPU checks of the great-circle distance the sphere PSE kernel uses in FP32:
theta^2 = u g(u), u = (chord / R)^2, g the Taylor series of 4 asin^2(s/2) / s^2
through s^12 (pair_kernels.hpp PseSphere), against the reference's
atan2(|x x y|, x . y) (SphereGeometry.f90:107-125)."""
import numpy as np

A = [1.0, 1 / 12, 1 / 90, 1 / 560, 1 / 3150, 1 / 16632, 1 / 84084]


def _theta2_series(u):
    g = np.zeros_like(u)
    for a in reversed(A):
        g = g * u + a
    return u * g


def _pairs(thmax, n=20000, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3))
    x /= np.linalg.norm(x, axis=1)[:, None]
    t = rng.normal(size=(n, 3))
    t -= (t * x).sum(1)[:, None] * x
    t /= np.linalg.norm(t, axis=1)[:, None]
    th = rng.uniform(0, thmax, n)
    y = np.cos(th)[:, None] * x + np.sin(th)[:, None] * t
    return x, y


def test_series_matches_reference_distance():
    """Where the kernel uses the series (7 eps <= 0.56 R, i.e. theta <= 0.56 in the
    kernel's support), it agrees with the reference's atan2 distance to 4e-9."""
    x, y = _pairs(0.56)
    ref = np.arctan2(np.linalg.norm(np.cross(x, y), axis=1), (x * y).sum(1)) ** 2
    u = ((x - y) ** 2).sum(1)
    got = _theta2_series(u)
    rel = np.abs(got - ref) / np.maximum(ref, 1e-300)
    assert rel.max() < 4e-9, rel.max()


def test_series_bounds_beyond_support():
    """Farther out the truncated series stays between s^2 and the exact theta^2, so
    the kernel value there is at most that of the chord distance (underflow)."""
    x, y = _pairs(np.pi, seed=1)
    ref = np.arctan2(np.linalg.norm(np.cross(x, y), axis=1), (x * y).sum(1)) ** 2
    u = ((x - y) ** 2).sum(1)
    got = _theta2_series(u)
    assert np.all(got <= ref * (1 + 1e-12)) and np.all(got >= u * (1 - 1e-12))


def test_asin_form_matches_reference_distance():
    """The general form theta = 2 asin(chord / 2R) is the same angle as atan2."""
    x, y = _pairs(np.pi, seed=2)
    ref = np.arctan2(np.linalg.norm(np.cross(x, y), axis=1), (x * y).sum(1))
    th = 2 * np.arcsin(np.minimum(0.5 * np.linalg.norm(x - y, axis=1), 1.0))
    assert np.abs(th - ref).max() < 1e-7  # asin is ill-conditioned only next to theta = pi
    near = ref < 3.0
    assert np.abs(th - ref)[near].max() < 1e-13
