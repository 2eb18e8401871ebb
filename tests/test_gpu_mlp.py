"""The far-field surrogate network on the GPU against the FP64 CPU oracle:
FP32 forward and gradients (hand-written GEMM backward), FP16-MFMA forward,
and training by the GPU's own Adam."""
import numpy as np
import pytest

from lpmgpu.mlp import ACT_ISRU, ACT_TANH, Mlp

pytestmark = pytest.mark.gpu

NIN, WIDTH, NH, NOUT = 13, 128, 3, 3
ACTS = [ACT_TANH, ACT_ISRU]  # oracle activation ids 1, 2


def _data(n, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, NIN)).astype(np.float32)
    t = np.stack([np.sin(x[:, 0]) * x[:, 3], np.cos(x[:, 1] + x[:, 4]), x[:, 2] * x[:, 5] * 0.3], 1)
    return x, t.astype(np.float32)


@pytest.mark.parametrize("act", ACTS)
def test_forward_fp32_matches_oracle(ctx, oracle, act):
    net = Mlp(ctx, NIN, WIDTH, NH, NOUT, seed=3, act=act)
    p = net.params()
    x, _ = _data(300)
    ref = oracle.mlp_forward(p.astype(np.float64), x.astype(np.float64), NIN, WIDTH, NH, NOUT, act)
    got = net.forward(x)
    e = np.abs(got - ref).max() / np.abs(ref).max()
    print("fp32 forward rel err", e)
    assert e < 1e-5


@pytest.mark.parametrize("act", ACTS)
def test_forward_mfma_fp16_matches_oracle(ctx, oracle, act):
    net = Mlp(ctx, NIN, WIDTH, NH, NOUT, seed=4, act=act)
    p = net.params()
    x, _ = _data(1000, seed=1)
    ref = oracle.mlp_forward(p.astype(np.float64), x.astype(np.float64), NIN, WIDTH, NH, NOUT, act)
    got = net.forward(x, mfma=True)
    e = np.sqrt(((got - ref) ** 2).sum() / (ref**2).sum())
    print("fp16 mfma forward rel rms err", e)
    assert e < 3e-3  # FP16 inputs/weights/activations, FP32 accumulation (measured 6e-4 / 7e-4)


@pytest.mark.parametrize("act", ACTS)
@pytest.mark.parametrize("n", [257, 4099])
def test_gradient_matches_oracle(ctx, oracle, n, act):
    """Hand-written backward (FP32) vs the FP64 oracle's analytic gradient, which
    is itself checked against central finite differences in the CPU tests.
    n = 4099 takes the split-K path of the weight gradients (ragged last split)."""
    net = Mlp(ctx, NIN, WIDTH, NH, NOUT, seed=5, act=act)
    p = net.params()
    x, t = _data(n, seed=2)
    loss, g = net.loss_grad(x, t)
    rl, rg = oracle.mlp_loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), NIN, WIDTH, NH,
                                  NOUT, act)
    e = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    print("loss", loss, rl, "grad rel err", e)
    assert abs(loss - rl) < 1e-5 * abs(rl)
    assert e < 1e-4


def test_training_reduces_loss(ctx):
    net = Mlp(ctx, NIN, WIDTH, NH, NOUT, seed=6)
    x, t = _data(20000, seed=3)
    l0, _ = net.loss_grad(x[:4096], t[:4096])
    net.train(x, t, epochs=20, batch=1024, lr=2e-3, seed=1)
    l1, _ = net.loss_grad(x[:4096], t[:4096])
    print("loss before / after", l0, l1)
    assert l1 < 0.2 * l0
    # the packed FP16 copy follows the trained FP32 weights
    y32 = net.forward(x[:2048])
    y16 = net.forward(x[:2048], mfma=True)
    assert np.sqrt(((y16 - y32) ** 2).sum() / (y32**2).sum()) < 3e-2
