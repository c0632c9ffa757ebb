"""GPU tests of the warm-start network kernels (BF16 matrix-core implicit
GEMMs, FP32 accumulation) against the float64 numpy oracle.

Tolerances are those of BF16 storage: activations and gradients are
rounded to 8 significant bits between layers (relative ~4e-3 each), so the
network output is checked to 3e-2 and every parameter gradient to 8e-2
relative L2."""
import numpy as np
import pytest

import i3elvis_amd as I
from oracle import net_oracle as NO

pytestmark = pytest.mark.gpu


def bf16(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _setup(n, seed):
    net = I.WarmStartNet(n, n, n, seed=seed)
    rng = np.random.default_rng(seed)
    p = net.get_params()
    # nonzero biases so that they are exercised
    ps = I.unflatten_params(p.copy())
    flat = []
    for i, a in enumerate(ps):
        if i % 2 == 1:
            a = rng.standard_normal(a.shape).astype(np.float32) * 0.05
        flat.append(np.asarray(a, np.float32).reshape(-1))
    p = np.concatenate(flat)
    net.set_params(p)
    x = rng.standard_normal((n, n, n, 2)).astype(np.float32)
    t = rng.standard_normal((n, n, n, 4)).astype(np.float32)
    return net, p, x, t


def _oracle_params(p):
    ps = I.unflatten_params(p)
    return [bf16(a) if i % 2 == 0 else a.astype(np.float64) for i, a in enumerate(ps)]


@pytest.mark.parametrize("n", [9, 17])
def test_forward_matches_oracle(gpu, n):
    print("case: net forward", n, flush=True)
    net, p, x, _ = _setup(n, 3)
    y = net.infer(x.reshape(-1, 2)).reshape(n, n, n, 4)
    yo, _ = NO.forward(_oracle_params(p), bf16(x))
    rel = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    print("forward rel L2", rel, flush=True)
    assert rel < 3e-2
    net.close()


def test_gradients_match_oracle(gpu):
    n = 17
    print("case: net gradient", n, flush=True)
    net, p, x, t = _setup(n, 4)
    loss, g = net.grad(x.reshape(-1, 2), t.reshape(-1, 4))
    ps = _oracle_params(p)
    yo, cache = NO.forward(ps, bf16(x))
    lo, gy = NO.loss_grad(yo, t.astype(np.float64))
    go = NO.backward(ps, cache, gy)
    assert abs(loss - lo) / lo < 3e-2
    gd = I.unflatten_params(g)
    for i, (a, b) in enumerate(zip(gd, go)):
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print("param", i, "rel L2", rel, flush=True)
        assert rel < 8e-2, (i, rel)
    net.close()


def test_adam_steps_reduce_loss(gpu):
    n = 17
    net, p, x, t = _setup(n, 5)
    # a smooth, learnable target: a linear map of the inputs
    t = np.stack([x[..., 0], -x[..., 1], x[..., 0] + x[..., 1], 0.5 * x[..., 0]], -1).astype(np.float32)
    l0 = net.train_step(x.reshape(1, -1, 2), t.reshape(1, -1, 4), lr=1e-3)
    for _ in range(30):
        l1 = net.train_step(x.reshape(1, -1, 2), t.reshape(1, -1, 4), lr=1e-3)
    print("loss", l0, "->", l1, flush=True)
    assert l1 < 0.5 * l0
    net.close()


def test_first_adam_step_matches_oracle(gpu):
    """one step from the same parameters and gradient: p - lr*sign(g)
    (Adam's first step), up to the gradient's BF16 noise on near-zero
    entries"""
    n = 9
    net, p, x, t = _setup(n, 6)
    _, g = net.grad(x.reshape(-1, 2), t.reshape(-1, 4))
    net.set_params(p)
    net.train_step(x.reshape(1, -1, 2), t.reshape(1, -1, 4), lr=1e-3)
    p1 = net.get_params()
    ps = [p.astype(np.float64)]
    st = dict(t=0, m=[np.zeros_like(ps[0])], v=[np.zeros_like(ps[0])])
    NO.adam(ps, [g.astype(np.float64)], st, lr=1e-3)
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    assert np.allclose(p1[big], ps[0][big], atol=2e-6)
    net.close()
