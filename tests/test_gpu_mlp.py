"""The learned scheduler's MLP kernels (BF16 weights, FP32 accumulation) against an
independent numpy fp64 reference on the same bf16-rounded weights: forward values,
backward gradients (also pinned by finite differences in test_np_ref.py), the Adam step,
and that training fits a known function."""
import ctypes as C

import numpy as np
import pytest

import np_ref as R
from gowalker_amd import engine as E
from gowalker_amd._abi import check, lib

pytestmark = pytest.mark.gpu

NIN, H1, H2 = 12, 64, 64


class Mlp:
    def __init__(self, ctx, seed=3):
        self.ctx = ctx
        self.h = C.c_void_p()
        check(lib().gw_mlp_create(ctx.h, NIN, H1, H2, seed, C.byref(self.h)), ctx.h)
        n = C.c_uint32()
        check(lib().gw_mlp_num_params(self.h, C.byref(n)))
        self.n = n.value

    def params(self):
        p = np.zeros(self.n, dtype=np.float32)
        check(lib().gw_mlp_get_params(self.ctx.h, self.h, p.ctypes.data), self.ctx.h)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().gw_mlp_set_params(self.ctx.h, self.h, p.ctypes.data), self.ctx.h)

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(len(x), dtype=np.float32)
        check(lib().gw_mlp_forward(self.ctx.h, self.h, x.ctypes.data, len(x), y.ctypes.data), self.ctx.h)
        return y

    def backward(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.n, dtype=np.float32)
        loss = C.c_float()
        check(lib().gw_mlp_backward(self.ctx.h, self.h, x.ctypes.data, t.ctypes.data, len(x), g.ctypes.data,
                                    C.byref(loss)), self.ctx.h)
        return g, loss.value

    def step(self, x, t, lr):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        check(lib().gw_mlp_step(self.ctx.h, self.h, x.ctypes.data, t.ctypes.data, len(x), lr, C.byref(loss)),
              self.ctx.h)
        return loss.value

    def close(self):
        lib().gw_mlp_destroy(self.ctx.h, self.h)


def test_params_are_bf16_and_roundtrip(gpu):
    with E.Context() as ctx:
        m = Mlp(ctx)
        assert m.n == NIN * H1 + H1 + H1 * H2 + H2 + H2 + 1
        p = m.params()
        np.testing.assert_array_equal(p, R.bf16_round(p))
        q = np.random.default_rng(1).normal(0, 0.3, m.n).astype(np.float32)
        m.set_params(q)
        np.testing.assert_array_equal(m.params(), R.bf16_round(q))
        m.close()


def test_forward_and_backward_match_reference(gpu):
    rng = np.random.default_rng(5)
    with E.Context() as ctx:
        m = Mlp(ctx)
        p = R.bf16_round(rng.normal(0, 0.3, m.n).astype(np.float32))
        m.set_params(p)
        x = rng.normal(0, 1, (777, NIN)).astype(np.float32)
        t = rng.normal(0, 1, 777).astype(np.float32)
        y = m.forward(x)
        yr, _ = R.mlp_forward(p.astype(np.float64), x.astype(np.float64), NIN, H1, H2)
        # fp32 accumulation over at most 64 terms: 1e-5 relative to the activation scale
        np.testing.assert_allclose(y, yr, rtol=1e-4, atol=1e-4 * np.abs(yr).max())
        g, loss = m.backward(x, t)
        gr, lr = R.mlp_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), NIN, H1, H2)
        assert abs(loss - lr) < 1e-4 * max(1.0, lr)
        np.testing.assert_allclose(g, gr, rtol=2e-3, atol=2e-5 * np.abs(gr).max())
        m.close()


def test_adam_step_reduces_loss_and_fits(gpu):
    rng = np.random.default_rng(9)
    with E.Context() as ctx:
        m = Mlp(ctx, seed=11)
        x = rng.normal(0, 1, (2048, NIN)).astype(np.float32)
        t = (np.log1p(np.abs(x[:, 0]) * 3) + 0.5 * x[:, 3] - 0.25 * x[:, 7] ** 2).astype(np.float32)
        first = m.step(x, t, 3e-3)
        for _ in range(400):
            last = m.step(x, t, 3e-3)
        assert last < 0.15 * first
        p = m.params()
        np.testing.assert_array_equal(p, R.bf16_round(p))   # working weights stay bf16
        m.close()
