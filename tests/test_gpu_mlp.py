"""The surrogate network's GPU kernels (dfargo_mlp_*: BF16 MFMA products with
FP32 accumulation, GELU, MSE, Adam) against the numpy restatement
tests/mlp_ref.py.  Bars, measured against the reference that rounds the
same operands to BF16: outputs within 1e-5 of their largest magnitude,
each layer's weight and bias gradients within 1e-4 (relative L2); against
plain FP64 (no BF16 rounding): 2e-2 both.  Measured on MI355X: 4e-7 and
3e-5 against the BF16-rounding reference, 5e-3 and 4e-3 against FP64.
Training on the device reduces a smooth regression loss tenfold."""
import numpy as np
import pytest

from dustyfargoadsg_amd.surrogate import MLP
from tests import mlp_ref

pytestmark = pytest.mark.gpu

N_IN, WIDTH, N_HIDDEN, N_OUT = 4, 256, 3, 256


def _data(B, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (B, N_IN)).astype(np.float32)
    r = np.linspace(0.4, 2.5, N_OUT)
    depth = 0.5 + 0.4 * x[:, :1]
    width = 0.1 + 0.05 * (x[:, 1:2] + 1)
    t = (1.0 - depth * np.exp(-((r[None, :] - 1.0) / width) ** 2) + 0.1 * x[:, 2:3] * (r[None, :] - 1.45))
    return x, t.astype(np.float32)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("B", [64, 200])
def test_forward_matches_reference(B):
    print(f"case mlp forward B={B}", flush=True)
    net = MLP(max_batch=256, seed=5)
    x, _ = _data(B, 1)
    y = net.forward(x)
    Ws, bs = mlp_ref.unflatten(net.params(), N_IN, WIDTH, N_HIDDEN, N_OUT)
    ye, _, _ = mlp_ref.forward(Ws, bs, x, emulate=True)
    yp, _, _ = mlp_ref.forward(Ws, bs, x, emulate=False)
    e1 = np.abs(y - ye).max() / np.abs(ye).max()
    e2 = np.abs(y - yp).max() / np.abs(yp).max()
    print("forward errors (bf16-emulating, fp64)", e1, e2, flush=True)
    assert e1 <= 1e-5 and e2 <= 2e-2


def test_gradients_match_reference():
    print("case mlp gradients B=96", flush=True)
    net = MLP(max_batch=256, seed=7)
    x, t = _data(96, 2)
    loss, g = net.gradients(x, t)
    p = net.params()
    Ws, bs = mlp_ref.unflatten(p, N_IN, WIDTH, N_HIDDEN, N_OUT)
    le, gWe, gbe = mlp_ref.loss_and_grads(Ws, bs, x, t, emulate=True)
    lp, gWp, gbp = mlp_ref.loss_and_grads(Ws, bs, x, t, emulate=False)
    assert abs(loss - le) <= 1e-5 * le
    gW, gb = mlp_ref.unflatten(g, N_IN, WIDTH, N_HIDDEN, N_OUT)
    for l in range(N_HIDDEN + 1):
        ew, eb = _rel(gW[l], gWe[l]), _rel(gb[l], gbe[l])
        pw, pb = _rel(gW[l], gWp[l]), _rel(gb[l], gbp[l])
        print(f"layer {l}: W {ew:.2e} b {eb:.2e} (fp64: {pw:.2e} {pb:.2e})", flush=True)
        assert ew <= 1e-4 and eb <= 1e-4, l
        assert pw <= 2e-2 and pb <= 2e-2, l


def test_training_reduces_the_loss_and_params_round_trip():
    print("case mlp training B=256", flush=True)
    net = MLP(max_batch=256, seed=9)
    x, t = _data(256, 3)
    l0, _ = net.gradients(x, t)
    l1 = net.train(x, t, 400, 1e-3)
    print("loss", l0, "->", l1, flush=True)
    assert l1 < 0.1 * l0
    p = net.params()
    other = MLP(max_batch=256, seed=1)
    other.set_params(p)
    assert np.array_equal(other.params(), p)
    assert np.array_equal(other.forward(x[:32]), net.forward(x[:32]))
