"""Pins the numpy references the GPU tests use: vectorised Philox against the oracle's,
and the numpy MLP gradient against central finite differences."""
import numpy as np

import _oracle as O
import np_ref as R


def test_numpy_philox_matches_oracle():
    rng = np.random.default_rng(3)
    c = rng.integers(0, 2**32, (4, 200), dtype=np.uint64).astype(np.uint32)
    k = [int(x) for x in rng.integers(0, 2**32, 2)]
    out = np.stack(R.philox(c[0], c[1], c[2], c[3], k[0], k[1]))
    for i in range(200):
        assert list(out[:, i]) == list(O.philox(c[:, i], k))


def test_numpy_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    nin, h1, h2, n = 5, 7, 6, 40
    npar = h1 * nin + h1 + h2 * h1 + h2 + h2 + 1
    p = rng.normal(0, 0.5, npar)
    x = rng.normal(0, 1, (n, nin))
    t = rng.normal(0, 1, n)
    g, loss = R.mlp_grad(p, x, t, nin, h1, h2)
    eps = 1e-6
    for i in rng.choice(npar, 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp = np.mean((R.mlp_forward(pp, x, nin, h1, h2)[0] - t) ** 2)
        lm = np.mean((R.mlp_forward(pm, x, nin, h1, h2)[0] - t) ** 2)
        assert abs((lp - lm) / (2 * eps) - g[i]) < 1e-5 * max(1.0, abs(g[i]))


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1.00390625, 1.01171875, -2.5, 3.0e-39], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1.0); 1 + 3*2^-8 -> 1 + 2^-6
    assert list(R.bf16_round(x)[:3]) == [1.0, 1.0, 1.015625]
    assert R.bf16_round(x)[3] == -2.5
