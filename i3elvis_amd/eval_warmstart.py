"""Evaluate the committed warm-start network on held-out seeds.

    python -m i3elvis_amd.eval_warmstart --seeds 40 [--seed0 2000] [--n 129]

For each held-out configuration-1 field (seeds seed0.., never used in
training) four mixed-precision solves to rel 1e-8 are run from different
starts, and their V-cycle counts compared:

  zero          the reference's start (no guess)
  lithostatic   the column integral of rho g
  network       lithostatic + the network's dynamic pressure and velocities
                (coefficients 1: the raw prediction)
  combined      the product's learned start: lithostatic + network outputs
                scaled by the residual-optimal coefficients, zero fallback

Prints one JSON line: per-seed rows and the means.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import make_config, Stokes
from . import warmstart as W

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warmstart_c1_129.npz")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=129)
    ap.add_argument("--seeds", type=int, default=40)
    ap.add_argument("--seed0", type=int, default=2000)
    ap.add_argument("--weights", default=WEIGHTS)
    a = ap.parse_args()
    n = a.n
    ws = W.WarmStart.load(a.weights, n=n)
    rows = []
    z = np.zeros(n ** 3)
    t0 = time.time()
    for s in range(a.seed0, a.seed0 + a.seeds):
        cfg = make_config(1, n, n, n, seed=s)
        S = Stokes(cfg)
        S.set_coefficients()
        row = dict(seed=s)
        S.set_solution(z, z, z, z)
        row["zero"] = S.solve(precision=32, rtol=1e-8).vcycles
        lith = W.lithostatic(cfg)
        _, row["rel0_lithostatic"] = S.combine_start([lith, None, None, None], [])
        row["lithostatic"] = S.solve(precision=32, rtol=1e-8).vcycles
        y = ws.net.infer(W.features(cfg)).astype(np.float64)
        net = [lith + y[:, 3] * ws.pscale, y[:, 0] * ws.vscale, y[:, 1] * ws.vscale, y[:, 2] * ws.vscale]
        _, row["rel0_network"] = S.combine_start(net, [])
        row["network"] = S.solve(precision=32, rtol=1e-8).vcycles
        info = W.learned_start(S, cfg, ws)
        row["rel0_combined"] = min(info["rel_learned"], info["rel_fallback"])
        row["coef_P"], row["coef_v"] = info["coef_dP"], info["coef_v"]
        row["combined"] = S.solve(precision=32, rtol=1e-8).vcycles
        S.close()
        rows.append(row)
        print(f"seed {s}: zero {row['zero']} lithostatic {row['lithostatic']} network {row['network']} "
              f"combined {row['combined']} (r0 lith {row['rel0_lithostatic']:.3e} net {row['rel0_network']:.3e} "
              f"comb {row['rel0_combined']:.3e})", file=sys.stderr, flush=True)
    mean = {k: float(np.mean([r[k] for r in rows])) for k in ("zero", "lithostatic", "network", "combined")}
    print(json.dumps(dict(n=n, seeds=[a.seed0, a.seed0 + a.seeds - 1], weights=os.path.basename(a.weights),
                          mean_vcycles=mean, wall_s=round(time.time() - t0, 1), rows=rows)))


if __name__ == "__main__":
    main()
