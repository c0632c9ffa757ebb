"""CPU tests pinning the warm-start network oracle: convolutions against
scipy's correlate, stride 2 as subsampled stride 1, upsampling transpose by
the adjoint identity, and every parameter gradient by central differences."""
import numpy as np
import scipy.ndimage as ndi

from oracle import net_oracle as N


def test_conv_single_channel_matches_scipy():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((7, 6, 5, 1))
    w = rng.standard_normal((27, 1, 1))
    k = w[:, 0, 0].reshape(3, 3, 3)
    ref = ndi.correlate(x[..., 0], k, mode="constant", cval=0.0)
    out = N.conv(x, w, np.zeros(1), 1)[..., 0]
    assert np.allclose(out, ref, atol=1e-12)


def test_conv_stride2_is_subsampled_stride1():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((9, 9, 7, 3))
    w = rng.standard_normal((27, 4, 3))
    b = rng.standard_normal(4)
    full = N.conv(x, w, b, 1)
    assert np.allclose(N.conv(x, w, b, 2), full[::2, ::2, ::2], atol=1e-12)


def test_conv_backward_adjoint():
    rng = np.random.default_rng(2)
    for s in (1, 2):
        x = rng.standard_normal((9, 7, 9, 3))
        w = rng.standard_normal((27, 5, 3))
        y = N.conv(x, w, np.zeros(5), s)
        g = rng.standard_normal(y.shape)
        dx, dw, db = N.conv_bwd(x, w, s, g)
        # <conv(x), g> is linear in x and in w
        assert np.isclose((y * g).sum(), (dx * x).sum(), rtol=1e-10)
        assert np.isclose((y * g).sum(), (dw * w).sum(), rtol=1e-10)
        assert np.allclose(db, g.sum(axis=(0, 1, 2)))


def test_up_adjoint_and_linear_exact():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((5, 3, 4, 2))
    g = rng.standard_normal((9, 5, 7, 2))
    assert np.isclose((N.up(a) * g).sum(), (a * N.up_t(g)).sum(), rtol=1e-12)
    z, x, y = np.meshgrid(np.arange(5), np.arange(3), np.arange(4), indexing="ij")
    lin = (1.5 * z - 2 * x + 0.25 * y)[..., None]
    z2, x2, y2 = np.meshgrid(np.arange(9), np.arange(5), np.arange(7), indexing="ij")
    assert np.allclose(N.up(lin)[..., 0], 0.75 * z2 - x2 + 0.125 * y2)


def test_network_gradients_by_central_differences():
    rng = np.random.default_rng(4)
    ps = N.init(5)
    for i in range(1, len(ps), 2):
        ps[i] = rng.standard_normal(ps[i].shape) * 0.1
    x = rng.standard_normal((9, 9, 9, N.CIN))
    tgt = rng.standard_normal((9, 9, 9, N.COUT))
    y, c = N.forward(ps, x)
    L, gy = N.loss_grad(y, tgt)
    grads = N.backward(ps, c, gy)
    h = 1e-6
    for li in range(len(ps)):
        flat = ps[li].reshape(-1)
        for j in rng.choice(flat.size, size=3, replace=False):
            old = flat[j]
            flat[j] = old + h
            Lp = N.loss_grad(N.forward(ps, x)[0], tgt)[0]
            flat[j] = old - h
            Lm = N.loss_grad(N.forward(ps, x)[0], tgt)[0]
            flat[j] = old
            fd = (Lp - Lm) / (2 * h)
            an = grads[li].reshape(-1)[j]
            assert abs(fd - an) <= 1e-5 * max(1e-6, abs(fd), abs(an)) + 1e-9, (li, j, fd, an)


def test_adam_first_step_is_signed_lr():
    ps = [np.array([1.0, -2.0, 3.0])]
    st = dict(t=0, m=[np.zeros(3)], v=[np.zeros(3)])
    N.adam(ps, [np.array([0.5, -4.0, 0.0])], st, lr=0.1)
    assert np.allclose(ps[0], [0.9, -1.9, 3.0], atol=1e-6)
