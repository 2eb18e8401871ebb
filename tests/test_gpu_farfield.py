"""Learned far field on the GPU: the near/far split, the cell expansion, the
training samples and the judge/fallback, against the plain sum and an
independent finite-difference Taylor expansion in numpy."""
import numpy as np
import pytest

from helpers import rel_err_max, taylor_cell_sum
from lpmgpu import ic, mesh
from lpmgpu.farfield import EXPANSION, LEARNED, PLAIN, FarFieldBVE

pytestmark = pytest.mark.gpu


def _case(level=5, cell_level=3):
    m = mesh.icosahedral_mesh(level, cell_level=cell_level)
    zeta = ic.gaussian_vortex(m.x, m.y, m.z, ic.lat_lon_point(0.3, 0.0), 4 * np.pi, 16.0) + ic.rh4_vorticity(
        m.x, m.y, m.z)
    return m, zeta


def test_all_near_equals_plain(ctx):
    """near_radius = pi: every cell is summed directly (self pairs excluded
    exactly), so the result is the plain sum up to FP32 summation order."""
    m, zeta = _case(4, 2)
    f = FarFieldBVE(ctx, m, zeta, near_radius=np.pi, band_radius=np.pi)
    p = f.velocity(PLAIN)
    e = f.velocity(EXPANSION)
    print("all-near vs plain", rel_err_max(e, p))
    assert rel_err_max(e, p) < 1e-6
    f.delete()


def test_binning_is_nearest_generator(ctx):
    """Every particle sits in the cell whose generator (the set-up area-weighted
    centroid direction of the given cell's active sources) is nearest; binning
    moves only a few sources off their mesh face."""
    m, zeta = _case(5, 3)
    f = FarFieldBVE(ctx, m, zeta, near_radius=0.1, band_radius=0.45)
    got = f.binning()
    P = m.xyz()
    act = m.active == 1
    gen = np.zeros((f.ncell, 3))
    np.add.at(gen, f.cell[act], P[act] * m.area[act, None])
    gen /= np.linalg.norm(gen, axis=1)[:, None]
    dots = P @ gen.T
    want = np.argmax(dots, axis=1)
    top2 = np.sort(dots, axis=1)[:, -2:]
    clear = top2[:, 1] - top2[:, 0] > 1e-5
    assert np.array_equal(got[clear], want[clear])
    assert (got[act] != f.cell[act]).mean() < 0.05
    f.delete()


def test_expansion_error_bounded_over_time(ctx):
    """Cells are re-binned at every evaluation, so the flow cannot stretch them:
    the expansion error after 30 steps stays near its initial value."""
    m, zeta = _case(5, 3)
    f = FarFieldBVE(ctx, m, zeta, rotation_rate=2 * np.pi, near_radius=0.15, band_radius=0.45)
    errs = []
    for _ in range(4):
        errs.append(rel_err_max(f.velocity(EXPANSION), f.velocity(PLAIN)))
        f.steps(0.01, 10, EXPANSION, judge_every=0)
    print("expansion error every 10 steps", errs)
    assert max(errs) < 3 * errs[0]
    f.delete()


@pytest.mark.parametrize("order", [2, 3, 4])
def test_expansion_error_falls_with_cutoff(ctx, order):
    m, zeta = _case(5, 3)
    errs = []
    for r in (0.05, 0.15, 0.3, 0.6):
        f = FarFieldBVE(ctx, m, zeta, near_radius=r, band_radius=r, order=order)
        errs.append(rel_err_max(f.velocity(EXPANSION), f.velocity(PLAIN)))
        f.delete()
    print(f"order {order}: expansion error vs near radius", errs)
    assert errs[-1] < 1e-4 and all(a > b for a, b in zip(errs, errs[1:]))


def test_fourth_order_beats_second(ctx):
    """Up/down triangle pairs cancel most of the third-order term in the sum; the
    fourth order removes the even remainder that is left."""
    m, zeta = _case(5, 3)
    e = {}
    for order in (2, 4):
        f = FarFieldBVE(ctx, m, zeta, near_radius=0.1, band_radius=0.1, order=order)
        e[order] = rel_err_max(f.velocity(EXPANSION), f.velocity(PLAIN))
        f.delete()
    print("near 0.1: order 2", e[2], "order 4", e[4])
    assert e[4] < e[2] / 3


def _cell_frame(P, idx, area):
    """Area-weighted centroid, frame rows (e1 towards the farthest source, e3
    along c) and whether that farthest source is unambiguous."""
    c = (P[idx] * area[idx, None]).sum(0) / area[idx].sum()
    d2 = ((P[idx] - c) ** 2).sum(1)
    top = np.sort(d2)[::-1]
    clear = len(top) < 2 or top[0] - top[1] > 1e-9 * top[0]
    a = P[idx[np.argmax(d2)]] - c
    e3 = c / np.linalg.norm(c)
    e1 = a - (a @ e3) * e3
    e1 /= np.linalg.norm(e1)
    return c, np.stack([e1, np.cross(e3, e1), e3]), clear


@pytest.mark.parametrize("order", [2, 4])
def test_training_samples_match_numpy(ctx, order):
    """A sample's inputs are the target and the cell's moments in the cell frame;
    its target is Q (exact cell sum - order-P expansion) l^2 r^(P+2) / W.  Check
    both with numpy: exact sum over the cell, expansion from Cauchy's integral
    for the Taylor coefficients of the cell sum (no moment formula involved)."""
    m, zeta = _case(5, 3)
    f = FarFieldBVE(ctx, m, zeta, near_radius=0.1, band_radius=0.45, order=order)
    f.collect(3000, seed=3)
    X, T, ids = f.samples(0, 3000)
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(T))
    assert (ids[:, 0] >= 0).mean() > 0.99
    cells = f.binning()
    act = m.active == 1
    w = zeta * m.area
    P = m.xyz()
    checked = 0
    for q in range(0, 3000, 23):
        p, s = ids[q]
        if p < 0:
            continue
        idx = np.flatnonzero(act & (cells == s))
        c, Q, clear = _cell_frame(P, idx, m.area)
        if not clear:  # symmetric cell: the anchor is a tie broken by rounding
            continue
        W = np.abs(w[idx]).sum()
        rho = np.max(np.arccos(np.clip(P[idx] @ Q[2], -1, 1)))
        ell = 2 * np.sin(0.5 * rho)
        x = P[p]
        d = P[idx] - c
        M0 = w[idx].sum()
        M1 = (w[idx, None] * d).sum(0)
        M2 = np.einsum("j,ja,jb->ab", w[idx], d, d)
        loc = Q @ (x - c) / ell
        M2l = Q @ M2 @ Q.T / (W * ell * ell)
        feat = np.concatenate([loc, [M0 / W], Q @ M1 / (W * ell), M2l[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]])
        assert np.allclose(X[q], feat, rtol=1e-5, atol=1e-5 * np.abs(feat).max()), (X[q], feat)

        exact = (w[idx, None] * P[idx] / (1.0 - P[idx] @ x)[:, None]).sum(0)
        expn = taylor_cell_sum(w[idx], P[idx], c, x, order)
        want = Q @ (exact - expn) * ell ** 2 * np.sqrt(loc @ loc) ** (order + 2) / W
        assert np.allclose(T[q], want, rtol=1e-3, atol=1e-4 * np.abs(want).max() + 1e-7), (T[q], want)
        checked += 1
    assert checked >= 60
    f.delete()


def test_network_gain_and_judge(ctx):
    """Train on samples from the method's own state, then compare the error of
    near + expansion (the method without the network) with near + expansion +
    network at the same cutoff; the judge falls back above tol."""
    m, zeta = _case(5, 3)
    f = FarFieldBVE(ctx, m, zeta, rotation_rate=2 * np.pi, near_radius=0.1, band_radius=0.45)
    f.collect(400000, seed=1)
    loss = f.train(epochs=20, batch=1024, lr=1e-3, seed=2)
    plain = f.velocity(PLAIN)
    e1 = rel_err_max(f.velocity(EXPANSION), plain)
    e2 = rel_err_max(f.velocity(LEARNED), plain)
    print(f"train loss {loss:.3e}; error without network {e1:.3e}, with network {e2:.3e}, gain {e1 / e2:.2f}x")
    assert np.isfinite(e2) and e2 < e1 / 1.5
    r = f.steps(0.01, 4, LEARNED, judge_every=2, tol=1e-12)
    assert r["mode"] == PLAIN and r["judged"] == 1 and r["last_error"] > 1e-12
    r = f.steps(0.01, 4, EXPANSION, judge_every=2, tol=1.0)
    assert r["mode"] == EXPANSION and r["judged"] == 2 and 0 <= r["last_error"] < 1.0
    st = f.state()
    assert np.abs(np.sqrt(st["x"] ** 2 + st["y"] ** 2 + st["z"] ** 2) - 1).max() < 1e-14
    f.delete()


def _geometry(P, area, cells, act, ncell):
    """Per cell, as the GPU computes it in FP64: direction and angular radius of
    the area-weighted source centroid, and of the (unweighted) target sum."""
    src = []
    for s in range(ncell):
        idx = np.flatnonzero(act & (cells == s))
        c = (P[idx] * area[idx, None]).sum(0) / area[idx].sum()
        u = c / np.linalg.norm(c)
        src.append((idx, c, u, np.arccos(np.clip(P[idx] @ u, -1, 1)).max()))
    tgt = []
    for t in range(ncell):
        idx = np.flatnonzero(cells == t)
        v = P[idx].sum(0)
        v /= np.linalg.norm(v)
        tgt.append((v, np.arccos(np.clip(P[idx] @ v, -1, 1)).max()))
    return src, tgt


@pytest.mark.parametrize("order", [2, 3, 4])
def test_coarse_level_matches_numpy(ctx, order):
    """lpm_bve_ff_set_parents: beyond the parent radius four sibling far cells
    are one expansion of their union about its own centroid.  Restated in numpy
    for 40 targets: the near cells summed directly (self excluded), every far
    record (a cell, or a parent where its four cells are far and the parent
    criterion holds) by its Taylor expansion from Cauchy's integral.  A parent
    radius no pair reaches gives the plain cell list, bitwise."""
    m, zeta = _case(5, 3)
    near, band, rpar = 0.1, 0.2, 0.3
    f = FarFieldBVE(ctx, m, zeta, near_radius=near, band_radius=band, order=order)
    e_cells = np.array(f.velocity(EXPANSION))
    n_cells = f.far_counts()
    f.set_parents(10.0)
    np.testing.assert_array_equal(np.array(f.velocity(EXPANSION)), e_cells)
    np.testing.assert_array_equal(f.far_counts(), n_cells)
    f.set_parents(rpar)
    got = np.array(f.velocity(EXPANSION))
    n_par = f.far_counts()
    cells = f.binning()
    plain = np.array(f.velocity(PLAIN))
    print(f"records per target cell: {n_cells.mean():.1f} cells only, {n_par.mean():.1f} with parents; error vs plain "
          f"{rel_err_max(e_cells, plain):.3e} / {rel_err_max(got, plain):.3e}", flush=True)
    assert n_par.mean() < 0.6 * n_cells.mean()
    assert rel_err_max(got, plain) < 3.0 * rel_err_max(e_cells, plain)

    ncell = f.ncell
    P, act, w = m.xyz(), m.active == 1, zeta * m.area
    src, tgt = _geometry(P, m.area, cells, act, ncell)
    par = []
    for g in range(ncell // 4):
        idx = np.concatenate([src[4 * g + k][0] for k in range(4)])
        c = (P[idx] * m.area[idx, None]).sum(0) / m.area[idx].sum()
        u = c / np.linalg.norm(c)
        par.append((idx, c, u, np.arccos(np.clip(P[idx] @ u, -1, 1)).max()))

    def ang(a, b):
        return np.arccos(np.clip(a @ b, -1, 1))

    scale = np.sqrt((plain ** 2).sum(0)).max()
    rng = np.random.default_rng(7)
    checked, n_parent_records = 0, 0
    for p in rng.choice(m.n, 60, replace=False):
        t = cells[p]
        v, rt = tgt[t]
        margin = np.inf
        V = np.zeros(3)
        x = P[p]
        records = 0
        for g in range(ncell // 4):
            cls = []
            for s in range(4 * g, 4 * g + 4):
                a = ang(v, src[s][2])
                pad = rt + src[s][3]
                margin = min(margin, abs(a - near - pad), abs(a - band - pad))
                cls.append(2 if a < near + pad else (1 if a < band + pad else 0))
            ap = ang(v, par[g][2])
            if all(k == 0 for k in cls):
                margin = min(margin, abs(ap - rpar - rt - par[g][3]))
            if all(k == 0 for k in cls) and ap >= rpar + rt + par[g][3]:
                idx, c = par[g][0], par[g][1]
                V += taylor_cell_sum(w[idx], P[idx], c, x, order)
                records += 1
                n_parent_records += 1
                continue
            for k, s in enumerate(range(4 * g, 4 * g + 4)):
                idx, c = src[s][0], src[s][1]
                if cls[k] == 2:
                    idx = idx[idx != p]
                    V += (w[idx, None] * P[idx] / (1.0 - P[idx] @ x)[:, None]).sum(0)
                else:
                    V += taylor_cell_sum(w[idx], P[idx], c, x, order)
                    records += 1
        if margin < 1e-9:  # a class decision within rounding of its threshold
            continue
        assert records == n_par[t]
        want = np.cross(V, x) / (4 * np.pi)
        assert np.abs(got[:, p] - want).max() < 2e-5 * scale, (p, got[:, p], want)
        checked += 1
        if checked == 40:
            break
    assert checked == 40 and n_parent_records > 0
    with pytest.raises(Exception):
        f.set_parents(-1.0)
    f.delete()


def test_coarse_level_needs_sibling_cells(ctx):
    """A cell count not divisible by 4 has no sibling groups: the coarse level
    is refused."""
    m = mesh.cubed_sphere_mesh(3, cell_level=0)
    zeta = ic.rh4_vorticity(m.x, m.y, m.z)
    f = FarFieldBVE(ctx, m, zeta, near_radius=0.1, band_radius=0.2)
    assert f.ncell == 6
    with pytest.raises(Exception):
        f.set_parents(0.6)
    f.set_parents(0.0)
    f.delete()


def test_online_fine_tuning_threshold(ctx):
    """lpm_bve_ff_set_online_from: a judge fine-tunes the network only when its
    error is above the given fraction of tol."""
    m, zeta = _case(4, 2)
    f = FarFieldBVE(ctx, m, zeta, rotation_rate=2 * np.pi, near_radius=0.1, band_radius=0.45)
    f.collect(20000, seed=1)
    f.train(epochs=2, batch=1024, lr=1e-3, seed=2)
    f.set_online(2000, epochs=1, batch=512, lr=1e-3, above=1.0)  # errors never exceed tol
    p0 = f.params()
    r = f.steps(0.01, 2, LEARNED, judge_every=1, tol=1.0)
    assert r["mode"] == LEARNED and r["judged"] == 2
    np.testing.assert_array_equal(f.params(), p0)
    f.set_online(2000, epochs=1, batch=512, lr=1e-3, above=0.0)
    f.steps(0.01, 1, LEARNED, judge_every=1, tol=1.0)
    assert not np.array_equal(f.params(), p0)
    with pytest.raises(Exception):
        f.set_online(2000, above=1.5)
    f.delete()
