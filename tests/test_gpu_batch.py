"""Batched sphere problems (one persistent launch per evaluation) against
per-problem solvers and the FP64 oracle."""
import numpy as np
import pytest

from helpers import rel_err_max
from lpmgpu import BVEBatch, BVESolver, workloads

pytestmark = pytest.mark.gpu


def _small(n_inst=6, seed=7):
    return workloads.cfg3(n_inst=n_inst, levels=(1, 4), seed=seed)["problems"]


@pytest.mark.parametrize("project", [False, True])
def test_batch_fp64_matches_single_solvers(ctx, project):
    probs = _small()
    b = BVEBatch(ctx, probs, precision="fp64", project=project)
    b.timestep(0.01, 2)
    got = b.split(b.state())
    for p, g in zip(probs, got):
        s = BVESolver(ctx, p["x"], p["y"], p["z"], p["relvort"], p["area"], p["active"],
                      rotation_rate=p["rotation_rate"], precision="fp64", project=project)
        s.timestep(0.01, 2)
        r = s.state()
        s.delete()
        assert rel_err_max((g["x"], g["y"], g["z"]), (r["x"], r["y"], r["z"])) < 1e-13
        assert rel_err_max((g["u"], g["v"], g["w"]), (r["u"], r["v"], r["w"])) < 1e-12
        assert rel_err_max(g["relvort"], r["relvort"]) < 1e-13
    b.delete()


def test_batch_fp32_matches_oracle(ctx, oracle):
    probs = _small(5, seed=11)
    b = BVEBatch(ctx, probs, precision="fp32")
    b.timestep(0.01, 2)
    got = b.split(b.state())
    for p, g in zip(probs, got):
        u = oracle.bve_sphere_velocity(p["x"], p["y"], p["z"], p["relvort"], p["area"], p["active"])
        st = dict(x=p["x"], y=p["y"], z=p["z"], relvort=p["relvort"], u=u[0], v=u[1], w=u[2])
        r = oracle.bve_timestep(st, p["area"], p["active"], 1.0, p["rotation_rate"], 0.01, nsteps=2)
        e = rel_err_max((g["u"], g["v"], g["w"]), (r["u"], r["v"], r["w"]))
        print("batch fp32 level", p["level"], "vel err", e)
        assert e < 1e-6
        assert rel_err_max((g["x"], g["y"], g["z"]), (r["x"], r["y"], r["z"])) < 1e-6
    b.delete()


def test_batch_graph_replay_matches_eager(ctx):
    probs = _small(4, seed=3)
    a = BVEBatch(ctx, probs, precision="fp32", project=True)
    g = BVEBatch(ctx, probs, precision="fp32", project=True, use_graph=True)
    for _ in range(3):
        a.timestep(0.01)
        g.timestep(0.01)
        sa, sg = a.state(), g.state()
        for k in sa:
            assert np.array_equal(sa[k], sg[k]), k
    a.delete()
    g.delete()


def test_batch_autotune_and_tile_shapes_agree(ctx):
    probs = _small(4, seed=5)
    b = BVEBatch(ctx, probs, precision="fp32")
    ref = b.state()
    best = b.autotune()
    print("autotune", best, b.tuning())
    assert best["targets_per_lane"] in (2, 4, 8) and best["ms"] > 0
    for shape in [(2, 128, 4096), (8, 256, 65536)]:
        b.set_tuning(*shape)
        c = BVEBatch(ctx, probs, precision="fp32")
        c.set_tuning(*shape)
        c.timestep(0.01)
        b2 = BVEBatch(ctx, probs, precision="fp32")
        b2.timestep(0.01)
        assert rel_err_max((c.state()["u"],), (b2.state()["u"],)) < 1e-6
        c.delete()
        b2.delete()
    b.delete()
    assert np.isfinite(ref["u"]).all()


def test_batch_tuning_rejects_bad_shapes(ctx):
    """Shapes without a kernel instance or with a chunk shorter than a tile
    are refused with an error, and the batch keeps its previous shape."""
    from lpmgpu.api import LpmError

    b = BVEBatch(ctx, _small(3, seed=11), precision="fp32")
    before = b.tuning()
    for shape in [(1, 256, 16384), (3, 256, 16384), (4, 100, 16384), (4, 0, 16384), (4, 256, 128)]:
        with pytest.raises(LpmError):
            b.set_tuning(*shape)
        assert b.tuning() == before
    b.timestep(0.01)
    assert np.isfinite(b.state()["u"]).all()
    b.delete()
