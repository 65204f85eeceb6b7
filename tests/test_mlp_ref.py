"""Pins the numpy restatement of the surrogate MLP (tests/mlp_ref.py) that
the GPU kernels are checked against: its analytic gradients equal central
finite differences of its own loss in FP64 (a dropped GELU derivative term,
a transposed product or a missing factor 2 in the loss gradient fails), and
the BF16 rounding is exact on representable values and round-to-nearest-even
on ties."""
import numpy as np

from tests import mlp_ref


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    n_in, width, n_hidden, n_out, B = 4, 8, 2, 6, 5
    sizes = [n_in] + [width] * n_hidden + [n_out]
    Ws = [rng.standard_normal((k, n)) * 0.5 for k, n in zip(sizes[:-1], sizes[1:])]
    bs = [rng.standard_normal(n) * 0.1 for n in sizes[1:]]
    x, t = rng.standard_normal((B, n_in)), rng.standard_normal((B, n_out))
    loss, gW, gb = mlp_ref.loss_and_grads(Ws, bs, x, t, emulate=False)
    p = mlp_ref.flatten(Ws, bs)
    g = mlp_ref.flatten(gW, gb)
    eps = 1e-6
    for i in rng.choice(p.size, 40, replace=False):
        q = p.copy()
        q[i] += eps
        lp = mlp_ref.loss_and_grads(*mlp_ref.unflatten(q, n_in, width, n_hidden, n_out), x, t, emulate=False)[0]
        q[i] -= 2 * eps
        lm = mlp_ref.loss_and_grads(*mlp_ref.unflatten(q, n_in, width, n_hidden, n_out), x, t, emulate=False)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0])
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 3 2^-8 to 1 + 2^-6
    assert list(mlp_ref.bf16(v)) == [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0]
