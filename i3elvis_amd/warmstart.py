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

    def __init__(self, n, seed=0):
        self.n = n
        self.net = WarmStartNet(n, n, n, seed=seed)
        self.vscale = 1.0
        self.pscale = 1.0

    def save(self, path):
        np.savez_compressed(path, params=self.net.get_params(), n=self.n, vscale=self.vscale, pscale=self.pscale)

    @classmethod
    def load(cls, path, n=None):
        """weights do not depend on the grid size: n (8k+1) may differ from
        the training grid"""
        d = np.load(path)
        w = cls(int(d["n"]) if n is None else n)
        w.net.set_params(d["params"])
        w.vscale, w.pscale = float(d["vscale"]), float(d["pscale"])
        return w

    def guess(self, cfg):
        y = self.net.infer(features(cfg))
        return solution_from(cfg, y, self.vscale, self.pscale)

    def train(self, data, epochs, batch=4, lr=1e-3, log=None, seed=0):
        """data: list of (features [N,2], targets [N,4]); returns losses"""
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
            if log:
                log(f"epoch {ep + 1}/{epochs} loss {losses[-1]:.4e}")
        return losses


def initial_rel(S, sol):
    """relative FP64 residual of a guess as the solver starts from it
    (missing unknowns zeroed, boundary rows applied); returns (rel, guess)"""
    S.set_solution(*sol)
    g = S.get_solution()
    nr = S.residual(*g)[1]
    return nr[0] / nr[1], g


def solve_with_guess(S, cfg, guess, fallback=None, **kw):
    """Start the solve from the better of the guess and the fallback
    (zero when None) by FP64 residual; returns (stats, which, rel0s)."""
    n = cfg["nx"] * cfg["ny"] * cfg["nz"]
    fb = fallback if fallback is not None else [np.zeros(n)] * 4
    relg, g = initial_rel(S, guess)
    relf, f = initial_rel(S, fb)
    use = g if relg < relf else f
    S.set_solution(*use)
    st = S.solve(**kw)
    return st, ("network" if relg < relf else "fallback"), (relg, relf)


def timed(f, *a, **k):
    t = time.time()
    r = f(*a, **k)
    return r, time.time() - t
