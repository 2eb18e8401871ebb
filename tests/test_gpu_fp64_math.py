"""The FP64 kernels' own reciprocal (hardware estimate + two Newton steps) and
logarithm (atanh series after an exponent split), checked term by term: with
a single active source every output is one pair term, which numpy computes
from the same separation with its correctly rounded division and log.  The
sums' parity with the oracle is tested elsewhere; here each term must agree to
a few ulp, so a wrong coefficient or range reduction cannot hide inside a sum."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sphere_points(rng, n):
    p = rng.normal(size=(3, n))
    p /= np.linalg.norm(p, axis=0)
    return p


def _one_source(n_targets, rng):
    """Particle 0 is the only active source; particles 1.. are the targets."""
    p = _sphere_points(rng, n_targets + 1)
    act = np.zeros(n_targets + 1, np.int32)
    act[0] = 1
    area = np.zeros(n_targets + 1)
    area[0] = 0.37
    q = np.zeros(n_targets + 1)
    q[0] = 1.9
    return p, act, area, q


def _on_sphere_solver(ctx, p, q, area, act):
    """A device-resident FP64 solver with projected stages: its sums use the
    on-sphere form 1/2 |x - y|^2 of R^2 - x.y, with no R^2 - |x|^2 terms, so
    numpy can form the same denominator from the solver's positions."""
    from lpmgpu import BVESolver

    x, y, z = p
    s = BVESolver(ctx, x, y, z, q, area, act, absvort=2 * q, precision="fp64", project=True)
    st = s.state()
    pr, pa = s.stream()
    s.delete()
    return st, pr, pa


def test_fp64_log_in_sphere_green_sum(ctx):
    rng = np.random.default_rng(7)
    p, act, area, q = _one_source(20000, rng)
    # 200 targets within 1e-6 ... 1e-3 of the source, the rest anywhere
    near = p[:, :1] + rng.normal(size=(3, 200)) * np.geomspace(1e-6, 1e-3, 200)
    p[:, 1:201] = near / np.linalg.norm(near, axis=0)
    st, pr, pa = _on_sphere_solver(ctx, p, q, area, act)
    P = np.array([st["x"], st["y"], st["z"]])
    d = P[:, 1:] - P[:, :1]
    dd = 0.5 * np.sum(d * d, axis=0)  # R^2 - x.y in the kernel's form
    want = -np.log(dd) / (4 * np.pi) * q[0] * area[0]
    got = pr[1:]
    # |log| > 1: an ulp of difference in the squared distance (fma vs separate
    # products) moves log by 2e-16 absolute, below the tolerance only there
    big = np.abs(np.log(dd)) > 1.0
    rel = np.abs(got - want)[big] / np.abs(want)[big]
    assert rel.max() < 2e-15, rel.max()
    assert np.abs(got - want).max() < 4e-16 * np.abs(want).max()
    assert np.array_equal(pa[1:], 2 * pr[1:])


def test_fp64_log_in_plane_green_sum(ctx):
    rng = np.random.default_rng(8)
    n = 20000
    x = np.concatenate([[0.1], rng.uniform(-3, 3, n)])
    y = np.concatenate([[-0.2], rng.uniform(-3, 3, n)])
    # a few targets very close to the source and very far from it
    x[1:50] = 0.1 + rng.uniform(-1e-7, 1e-7, 49)
    y[1:50] = -0.2 + rng.uniform(-1e-7, 1e-7, 49)
    x[50:100] = 0.1 + rng.uniform(-1e4, 1e4, 50)
    act = np.zeros(n + 1, np.int32)
    act[0] = 1
    area = np.zeros(n + 1)
    area[0] = 0.5
    w = np.zeros(n + 1)
    w[0] = -1.3
    psi = ctx.plane_stream(x, y, w, area, act, precision="fp64")
    r2 = (x[1:] - x[0]) ** 2 + (y[1:] - y[0]) ** 2
    want = 0.5 * np.log(r2) / (2 * np.pi) * w[0] * area[0]
    big = np.abs(np.log(r2)) > 1.0
    rel = np.abs(psi[1:] - want)[big] / np.abs(want)[big]
    assert rel.max() < 2e-15, rel.max()


def test_fp64_reciprocal_in_biot_savart_sum(ctx):
    rng = np.random.default_rng(9)
    p, act, area, q = _one_source(20000, rng)
    near = p[:, :1] + rng.normal(size=(3, 200)) * np.geomspace(1e-6, 1e-3, 200)
    p[:, 1:201] = near / np.linalg.norm(near, axis=0)
    st, _, _ = _on_sphere_solver(ctx, p, q, area, act)
    P = np.array([st["x"], st["y"], st["z"]])
    # u_i = s (x_i - y) x x_i = s (x_i x y),  s = -zeta A / (4 pi (R^2 - x_i . y)),
    # R^2 - x.y = |x - y|^2 / 2; error scale |s| |x_i - y| (the size of each factor)
    d = P[:, 1:] - P[:, :1]
    den = 0.5 * np.sum(d * d, axis=0)
    s = -q[0] * area[0] / (4 * np.pi * den)
    want = np.cross(d.T, P[:, 1:].T).T * s
    got = np.array([st["u"][1:], st["v"][1:], st["w"][1:]])
    err = np.linalg.norm(got - want, axis=0) / (np.abs(s) * np.linalg.norm(d, axis=0))
    assert err.max() < 1e-14, err.max()
