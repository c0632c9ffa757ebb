"""Warm-start network oracle — TEST INFRASTRUCTURE ONLY.

A numpy (float64) statement of the 3D convolutional encoder-decoder that the
device library trains and evaluates: forward pass, the gradients of every
parameter by the chain rule, and one Adam step.  Only tests/ use it, as the
checker of the device kernels; its own correctness is pinned by finite
differences and by library convolutions in tests/test_net_oracle.py.

Conventions (shared with the device code):
  activations  [D, H, W, C]   (z, x, y order of the node grid, C fastest)
  weights      [27, Cout, Cin] tap t = 9*tz + 3*tx + ty, offsets t* - 1
  convolution  out[o] = b + sum_t sum_ci w[t, :, ci] * in[o*s + t - 1, ci]
               (zero outside the grid); stride s in {1, 2}
  upsampling   node-linear 2n-1 <- n: even nodes copy, odd nodes average
Network (levels 16/32/64/64 channels, ReLU after every conv but the last):
  e1 = relu(conv(x, 1))          n
  e2 = relu(conv(e1, 2))         (n+1)/2
  e3 = relu(conv(e2, 2))         ...
  e4 = relu(conv(e3, 2))
  d3 = relu(conv(up(e4) + e3, 1))
  d2 = relu(conv(up(d3) + e2, 1))
  y  = conv(up(d2) + e1, 1)
"""
import numpy as np

CH = (16, 32, 64, 64)
CIN, COUT = 2, 4
# (cin, cout, stride) of the seven convolutions
LAYERS = [(CIN, CH[0], 1), (CH[0], CH[1], 2), (CH[1], CH[2], 2), (CH[2], CH[3], 2),
          (CH[3], CH[1], 1), (CH[1], CH[0], 1), (CH[0], COUT, 1)]


def _taps():
    return [(tz, tx, ty) for tz in range(3) for tx in range(3) for ty in range(3)]


def conv(x, w, b, s):
    """x [D,H,W,Ci], w [27,Co,Ci], b [Co] -> [Do,Ho,Wo,Co]"""
    D, H, W, _ = x.shape
    Do, Ho, Wo = (D - 1) // s + 1, (H - 1) // s + 1, (W - 1) // s + 1
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.broadcast_to(b, (Do, Ho, Wo, len(b))).astype(np.float64).copy()
    for t, (tz, tx, ty) in enumerate(_taps()):
        sl = xp[tz:tz + s * (Do - 1) + 1:s, tx:tx + s * (Ho - 1) + 1:s, ty:ty + s * (Wo - 1) + 1:s]
        out += sl @ w[t].T
    return out


def conv_bwd(x, w, s, g):
    """gradients of conv(x, w, b, s) given g = dL/dout: (dx, dw, db)"""
    D, H, W, _ = x.shape
    Do, Ho, Wo = g.shape[:3]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp, dtype=np.float64)
    dw = np.zeros_like(w, dtype=np.float64)
    for t, (tz, tx, ty) in enumerate(_taps()):
        idx = (slice(tz, tz + s * (Do - 1) + 1, s), slice(tx, tx + s * (Ho - 1) + 1, s),
               slice(ty, ty + s * (Wo - 1) + 1, s))
        dw[t] = np.einsum("zxyo,zxyi->oi", g, xp[idx])
        dxp[idx] += g @ w[t]
    db = g.sum(axis=(0, 1, 2))
    return dxp[1:-1, 1:-1, 1:-1], dw, db


def _up1(a, axis):
    n = a.shape[axis]
    shp = list(a.shape)
    shp[axis] = 2 * n - 1
    out = np.zeros(shp, dtype=np.float64)
    ev = [slice(None)] * a.ndim
    od = [slice(None)] * a.ndim
    ev[axis] = slice(0, None, 2)
    od[axis] = slice(1, None, 2)
    lo = [slice(None)] * a.ndim
    hi = [slice(None)] * a.ndim
    lo[axis] = slice(0, n - 1)
    hi[axis] = slice(1, n)
    out[tuple(ev)] = a
    out[tuple(od)] = 0.5 * (a[tuple(lo)] + a[tuple(hi)])
    return out


def up(a):
    return _up1(_up1(_up1(a, 0), 1), 2)


def _up1_t(g, axis):
    m = g.shape[axis]
    n = (m + 1) // 2
    ev = [slice(None)] * g.ndim
    od = [slice(None)] * g.ndim
    ev[axis] = slice(0, None, 2)
    od[axis] = slice(1, None, 2)
    out = g[tuple(ev)].astype(np.float64).copy()
    go = 0.5 * g[tuple(od)]
    lo = [slice(None)] * g.ndim
    hi = [slice(None)] * g.ndim
    lo[axis] = slice(0, n - 1)
    hi[axis] = slice(1, n)
    out[tuple(lo)] += go
    out[tuple(hi)] += go
    return out


def up_t(g):
    return _up1_t(_up1_t(_up1_t(g, 2), 1), 0)


def init(seed=0):
    """He-normal weights, zero biases (the device initialiser draws its own
    numbers; this one is for the oracle's tests)."""
    rng = np.random.default_rng(seed)
    ps = []
    for ci, co, _ in LAYERS:
        ps.append(rng.standard_normal((27, co, ci)) * np.sqrt(2.0 / (27 * ci)))
        ps.append(np.zeros(co))
    return ps


def forward(ps, x):
    """returns (y, cache)"""
    W = ps[0::2]
    B = ps[1::2]
    a = []
    h = x
    for i in range(4):
        h = np.maximum(conv(h, W[i], B[i], LAYERS[i][2]), 0)
        a.append(h)
    e1, e2, e3, e4 = a
    i3 = up(e4) + e3
    d3 = np.maximum(conv(i3, W[4], B[4], 1), 0)
    i2 = up(d3) + e2
    d2 = np.maximum(conv(i2, W[5], B[5], 1), 0)
    i1 = up(d2) + e1
    y = conv(i1, W[6], B[6], 1)
    return y, dict(x=x, e=a, i3=i3, d3=d3, i2=i2, d2=d2, i1=i1)


def loss_grad(y, target):
    """mean squared error over all entries and its gradient"""
    d = y - target
    return float((d * d).mean()), 2.0 * d / d.size


def backward(ps, c, gy):
    W = ps[0::2]
    grads = [None] * len(ps)
    e1, e2, e3, e4 = c["e"]
    gi1, grads[12], grads[13] = conv_bwd(c["i1"], W[6], 1, gy)
    ge1 = gi1.copy()
    gd2 = up_t(gi1) * (c["d2"] > 0)
    gi2, grads[10], grads[11] = conv_bwd(c["i2"], W[5], 1, gd2)
    ge2 = gi2.copy()
    gd3 = up_t(gi2) * (c["d3"] > 0)
    gi3, grads[8], grads[9] = conv_bwd(c["i3"], W[4], 1, gd3)
    ge3 = gi3.copy()
    ge4 = up_t(gi3)
    g = ge4
    for i, (ein, eout, ge_add) in enumerate(((e3, e4, None), (e2, e3, ge3), (e1, e2, ge2), (c["x"], e1, ge1))):
        li = 3 - i
        if ge_add is not None:
            g = g + ge_add
        gz = g * (eout > 0)
        g, grads[2 * li], grads[2 * li + 1] = conv_bwd(ein, W[li], LAYERS[li][2], gz)
    return grads


def adam(ps, grads, st, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """one Adam step in place; st = dict(t, m, v)"""
    st["t"] += 1
    t = st["t"]
    for i, (p, g) in enumerate(zip(ps, grads)):
        st["m"][i] = b1 * st["m"][i] + (1 - b1) * g
        st["v"][i] = b2 * st["v"][i] + (1 - b2) * g * g
        mh = st["m"][i] / (1 - b1 ** t)
        vh = st["v"][i] / (1 - b2 ** t)
        p -= lr * mh / (np.sqrt(vh) + eps)
