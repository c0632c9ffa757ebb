"""Learned warm start for the Stokes solve: host side.

The network and its training run on the device through the i3g_warmstart_*
entry points (include/i3g.h); this module builds the scaled input/target
fields, drives the training loop and applies the fallback rule.

  inputs   c0 = (log10 eta - 21) / 2,  c1 = (rho - rho0) / 50       at nodes
  targets  vx, vy, vz / vscale,  (P - rho0 g_y (y - y_gauge)) / pscale
           (the hydrostatic part of P is removed; the solver's pressure
           gauge fixes P at the node (pxinit, pyinit, pzinit))
  fallback the solve starts from whichever of {network guess, given
           guess} has the smaller FP64 residual (the given guess is the
           previous step's solution in a time loop, zero otherwise).
"""
import time

import numpy as np

from . import WarmStartNet, make_config, Stokes

RHO0 = 3300.0


def features(cfg):
    lg = (np.log10(np.maximum(cfg["nu"].astype(np.float64), 1.0)) - 21.0) / 2.0
    dr = (cfg["ro"].astype(np.float64) - RHO0) / 50.0
    return np.stack([lg, dr], -1).astype(np.float32)


def lithostatic(cfg, gauge=(2, 1, 2)):
    """P = integral of rho g_y dy down each (x, z) column at the pressure
    points (cell centres, m1, m2, m3 >= 1: the cell's four corner densities
    averaged), shifted to zero at the gauge cell"""
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    ro = cfg["ro"].astype(np.float64).reshape(nz, nx, ny)
    gy = cfg["gy"]
    # cell-centre density in the (x, z) plane at every node row m2
    rc = np.zeros_like(ro)
    rc[1:, 1:, :] = 0.25 * (ro[1:, 1:, :] + ro[:-1, 1:, :] + ro[1:, :-1, :] + ro[:-1, :-1, :])
    yc = np.r_[0.0, 0.5 * (gy[1:] + gy[:-1])]
    P = np.zeros_like(ro)
    # between pressure rows m2-1 and m2 (m2 >= 2) the node row m2-1 lies
    for m2 in range(2, ny):
        P[:, :, m2] = P[:, :, m2 - 1] + cfg["g"][1] * rc[:, :, m2 - 1] * (yc[m2] - yc[m2 - 1])
    P -= P[gauge[2], gauge[0], gauge[1]]
    return P.reshape(-1)


def hydrostatic(cfg, gauge=(2, 1, 2)):
    """rho0 g_y (y - y_gauge) at every node (y is the vertical axis)"""
    ny, nx, nz = cfg["ny"], cfg["nx"], cfg["nz"]
    gy = cfg["gy"]
    m2 = np.tile(np.arange(ny), nx * nz)
    # P lives at cell centres (m2 - 1/2): use the centre depth
    yc = np.where(m2 > 0, 0.5 * (gy[np.maximum(m2 - 1, 0)] + gy[m2]), 0.0)
    yg = 0.5 * (gy[gauge[1] - 1] + gy[gauge[1]])
    return RHO0 * cfg["g"][1] * (yc - yg)


def targets(cfg, sol, vscale, pscale):
    P, vx, vy, vz = sol
    return np.stack([vx / vscale, vy / vscale, vz / vscale, (P - hydrostatic(cfg)) / pscale], -1).astype(np.float32)


def solution_from(cfg, y, vscale, pscale):
    y = y.astype(np.float64)
    return [y[:, 3] * pscale + hydrostatic(cfg), y[:, 0] * vscale, y[:, 1] * vscale, y[:, 2] * vscale]


def label(n, seed, precision=64, rtol=1e-8):
    """configuration-1 field of the given seed and its converged solution"""
    cfg = make_config(1, n, n, n, seed=seed)
    S = Stokes(cfg)
    S.set_coefficients()
    z = np.zeros(n ** 3)
    S.set_solution(z, z, z, z)
    st = S.solve(precision=precision, rtol=rtol)
    sol = S.get_solution()
    S.close()
    return cfg, sol, st


class WarmStart:
    """A trained network plus its scales."""

    def __init__(self, n, seed=0, shape=None):
        """shape: (nz, nx, ny) node grid when not a cube of n"""
        self.n = n
        self.net = WarmStartNet(*(shape or (n, n, n)), seed=seed)
        self.vscale = 1.0
        self.pscale = 1.0

    def save(self, path):
        np.savez_compressed(path, params=self.net.get_params(), n=self.n, vscale=self.vscale, pscale=self.pscale)

    @classmethod
    def load(cls, path, n=None, shape=None):
        """weights do not depend on the grid size: n (8k+1) or shape
        (nz, nx, ny; each 8k+1) may differ from the training grid"""
        d = np.load(path)
        w = cls(int(d["n"]) if n is None else n, shape=shape)
        w.net.set_params(d["params"])
        w.vscale, w.pscale = float(d["vscale"]), float(d["pscale"])
        return w

    def guess(self, cfg):
        y = self.net.infer(features(cfg))
        return solution_from(cfg, y, self.vscale, self.pscale)

    def train(self, data, epochs, batch=4, lr=1e-3, log=None, seed=0, checkpoint=None):
        """data: list of (features [N,2], targets [N,4]); returns losses.
        checkpoint: path written after every epoch"""
        rng = np.random.default_rng(seed)
        losses = []
        for ep in range(epochs):
            order = rng.permutation(len(data))
            tot = 0.0
            for b0 in range(0, len(order), batch):
                idx = order[b0:b0 + batch]
                xs = np.stack([data[i][0] for i in idx])
                ts = np.stack([data[i][1] for i in idx])
                tot += self.net.train_step(xs, ts, lr=lr) * len(idx)
            losses.append(tot / len(data))
            if checkpoint:
                self.save(checkpoint)
            if log:
                log(f"epoch {ep + 1}/{epochs} loss {losses[-1]:.4e}")
        return losses


def learned_start(S, cfg, ws, fallback=None):
    """Set the solver's start from the network: the lithostatic pressure
    plus the network's dynamic pressure and velocities, each scaled by the
    coefficient that minimises the FP64 residual (computed on the device,
    i3g_stokes_combine_start).  The fallback (the previous step's solution;
    zero when None) is used instead when its residual is smaller.  Returns
    info (coefficients, initial residuals, which start was taken)."""
    fb = fallback if fallback is not None else [None] * 4
    _, relf = S.combine_start(fb, [])
    y = ws.net.infer(features(cfg)).astype(np.float64)
    base = [lithostatic(cfg), None, None, None]
    dP = [y[:, 3] * ws.pscale, None, None, None]
    dv = [None, y[:, 0] * ws.vscale, y[:, 1] * ws.vscale, y[:, 2] * ws.vscale]
    _, rel_lith = S.combine_start(base, [])
    a, rel = S.combine_start(base, [dP, dv])
    start = "network"
    if relf < rel:
        S.combine_start(fb, [])
        start = "fallback"
    return dict(coef_dP=float(a[0]), coef_v=float(a[1]), rel_learned=rel, rel_lithostatic=rel_lith,
                rel_fallback=relf, start=start)


def learned_start_device(S, ws, gauge=(2, 1, 2)):
    """learned_start with the zero fallback, built entirely on the device
    from the context's own fields (i3g_stokes_learned_start): no host
    arrays and no host-device copies of fields"""
    return S.learned_start(ws.net, ws.pscale, ws.vscale, gauge)


def timed(f, *a, **k):
    t = time.time()
    r = f(*a, **k)
    return r, time.time() - t
