"""The learned scheduler's MLP kernels (BF16 weights, FP32 accumulation) against an
independent numpy fp64 reference on the same bf16-rounded weights: forward values,
backward gradients (also pinned by finite differences in test_np_ref.py), one Adam step
against a numpy fp32 restatement, determinism, and that training fits a known function.
The 12-64-64 scorer runs the fused single-workgroup kernel; 16-128-128 does not fit its
LDS budget and runs the multi-kernel path."""
import ctypes as C

import numpy as np
import pytest

import np_ref as R
from gowalker_amd import engine as E
from gowalker_amd._abi import check, lib

pytestmark = pytest.mark.gpu

NIN, H1, H2 = 12, 64, 64


class Mlp:
    def __init__(self, ctx, seed=3, shape=(NIN, H1, H2)):
        self.ctx = ctx
        self.h = C.c_void_p()
        self.shape = shape
        check(lib().gw_mlp_create(ctx.h, shape[0], shape[1], shape[2], seed, C.byref(self.h)), ctx.h)
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

    def step(self, x, t, lr, wait=True):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        check(lib().gw_mlp_step(self.ctx.h, self.h, x.ctypes.data, t.ctypes.data, len(x), lr,
                                C.byref(loss) if wait else None), self.ctx.h)
        return loss.value if wait else None

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


@pytest.mark.parametrize("shape", [(12, 64, 64), (16, 128, 128)])
def test_forward_and_backward_match_reference(gpu, shape):
    rng = np.random.default_rng(5)
    nin, h1, h2 = shape
    scale = 0.3 * (64 / h1) ** 0.5
    with E.Context() as ctx:
        m = Mlp(ctx, shape=shape)
        p = R.bf16_round(rng.normal(0, scale, m.n).astype(np.float32))
        m.set_params(p)
        x = rng.normal(0, 1, (777, nin)).astype(np.float32)
        t = rng.normal(0, 1, 777).astype(np.float32)
        y = m.forward(x)
        yr, _ = R.mlp_forward(p.astype(np.float64), x.astype(np.float64), nin, h1, h2)
        # fp32 accumulation over at most 128 terms: 1e-5 relative to the activation scale
        np.testing.assert_allclose(y, yr, rtol=1e-4, atol=1e-4 * np.abs(yr).max())
        g, loss = m.backward(x, t)
        gr, lr = R.mlp_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), nin, h1, h2)
        assert abs(loss - lr) < 1e-4 * max(1.0, lr)
        np.testing.assert_allclose(g, gr, rtol=2e-3, atol=2e-5 * np.abs(gr).max())
        # fixed summation order: a second call gives the same bits
        g2, loss2 = m.backward(x, t)
        np.testing.assert_array_equal(g, g2)
        assert loss == loss2
        m.close()


def test_adam_step_matches_restatement(gpu):
    rng = np.random.default_rng(13)
    with E.Context() as ctx:
        m = Mlp(ctx, seed=2)
        p = m.params()
        x = rng.normal(0, 1, (300, NIN)).astype(np.float32)
        t = rng.normal(0, 1, 300).astype(np.float32)
        g, _ = m.backward(x, t)
        lr = np.float32(1e-3)
        m.step(x, t, float(lr))
        f = np.float32
        a = (f(1) - f(0.9)) * g
        b = (f(1) - f(0.999)) * g * g
        c1, c2 = f(1) - f(0.9), f(1) - f(0.999)
        want = R.bf16_round((p - lr * (a / c1) / (np.sqrt(b / c2) + f(1e-8))).astype(np.float32))
        got = m.params()
        # within one bf16 ulp (fused multiply-adds may move the fp32 value by an ulp)
        ulp = np.abs(want) * 2.0 ** -7 + 1e-30
        assert np.all(np.abs(got - want) <= ulp)
        assert np.mean(got == want) > 0.99
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


@pytest.mark.parametrize("shape", [(12, 64, 64), (16, 128, 128)])
def test_queued_steps_equal_waited_steps(gpu, shape):
    """Adam steps queued without a host wait (loss NULL) leave the same parameters as steps
    that each wait for their loss: the batches' inputs go through the pinned ring in order."""
    rng = np.random.default_rng(8)
    batches = [(rng.normal(0, 1, (96, shape[0])).astype(np.float32), rng.normal(0, 1, 96).astype(np.float32))
               for _ in range(9)]
    out = []
    for wait in (True, False):
        with E.Context() as ctx:
            m = Mlp(ctx, seed=5, shape=shape)
            for x, t in batches:
                m.step(x, t, 1e-2, wait=wait)
            out.append(m.params())   # reads the parameters after the queued steps
            m.close()
    np.testing.assert_array_equal(out[0], out[1])
