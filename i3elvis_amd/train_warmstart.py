"""Train the warm-start network on the device and evaluate it.

    python -m i3elvis_amd.train_warmstart --n 129 --train 48 --eval 6 --epochs 30 --out weights.npz

Training fields are configuration-1 fields (seeds 0..train-1) labelled by
the plain FP64 multigrid to 1e-8; evaluation fields use the held-out seeds
2000.. and compare the mixed-precision solve from a zero guess with the
solve from the network's guess (fallback rule applied).  Prints one JSON
summary line at the end.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import make_config, Stokes
from . import warmstart as W


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=129)
    ap.add_argument("--train", type=int, default=48)
    ap.add_argument("--eval", type=int, default=6)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default="")
    ap.add_argument("--eval-seed0", type=int, default=2000)
    a = ap.parse_args()
    n = a.n
    t0 = time.time()
    sols, cfgs = [], []
    for s in range(a.train):
        cfg, sol, st = W.label(n, s, precision=64)
        log(f"label seed {s}: {st.vcycles} V-cycles {st.seconds:.2f}s rel {st.rel_res:.2e}")
        cfgs.append(cfg)
        sols.append(sol)
    tl = time.time() - t0
    v = np.concatenate([np.concatenate(s[1:]) for s in sols])
    vscale = float(np.sqrt(np.mean(v * v)))
    pp = np.concatenate([s[0] - W.hydrostatic(c) for s, c in zip(sols, cfgs)])
    pscale = float(np.sqrt(np.mean(pp * pp)))
    log(f"scales: v {vscale:.3e} m/s, P' {pscale:.3e} Pa")
    data = [(W.features(c), W.targets(c, s, vscale, pscale)) for c, s in zip(cfgs, sols)]
    del sols, cfgs
    ws = W.WarmStart(n, seed=1)
    ws.vscale, ws.pscale = vscale, pscale
    t1 = time.time()
    losses = ws.train(data, a.epochs, batch=a.batch, lr=a.lr, log=log, checkpoint=a.out or None)
    tt = time.time() - t1
    if a.out:
        ws.save(a.out)
    # evaluation on held-out seeds, mixed precision (the product path)
    rows = []
    for s in range(a.eval_seed0, a.eval_seed0 + a.eval):
        cfg = make_config(1, n, n, n, seed=s)
        S = Stokes(cfg)
        S.set_coefficients()
        z = np.zeros(n ** 3)
        S.set_solution(z, z, z, z)
        st0 = S.solve(precision=32, rtol=1e-8)
        ref = S.get_solution()
        graw = ws.guess(cfg)
        gerr = [float(np.linalg.norm(a1 - b1) / max(np.linalg.norm(b1), 1e-300)) for a1, b1 in zip(graw, ref)]
        t = time.time()
        info = W.learned_start(S, cfg, ws)
        tg = time.time() - t
        which, rg, rf = info["start"], info["rel_learned"], info["rel_fallback"]
        st1 = S.solve(precision=32, rtol=1e-8)
        # the lithostatic pressure alone, for comparison
        plith = W.lithostatic(cfg)
        S.set_solution(plith, z, z, z)
        st2 = S.solve(precision=32, rtol=1e-8)
        sol1 = S.get_solution()
        err = max(np.linalg.norm(a1 - b1) / max(np.linalg.norm(b1), 1e-300) for a1, b1 in zip(sol1[1:], ref[1:]))
        rows.append(dict(seed=s, zero_vcycles=st0.vcycles, zero_s=round(st0.seconds, 3), net_vcycles=st1.vcycles,
                         net_s=round(st1.seconds, 3), guess_s=round(tg, 3), start=which, rel0_net=rg, rel0_zero=rf,
                         lith_vcycles=st2.vcycles, lith_s=round(st2.seconds, 3), guess=info,
                         raw_guess_rel_err_P_vx_vy_vz=gerr, vel_rel_diff=float(err)))
        log(f"eval seed {s}: zero {st0.vcycles} V-cycles {st0.seconds:.2f}s | lithostatic {st2.vcycles} V-cycles "
            f"(r0 {info['rel_lithostatic']:.2e}) | learned ({which}, r0 {rg:.2e}, coef P {info['coef_dP']:.2f} "
            f"v {info['coef_v']:.2f}; raw error P/vx/vy/vz {' '.join(f'{e:.2f}' for e in gerr)}) {st1.vcycles} "
            f"V-cycles {st1.seconds:.2f}s + guess {tg:.2f}s")
        S.close()
    zv = np.mean([r["zero_vcycles"] for r in rows])
    nv = np.mean([r["net_vcycles"] for r in rows])
    lv = np.mean([r["lith_vcycles"] for r in rows])
    zs = np.mean([r["zero_s"] for r in rows])
    ns = np.mean([r["net_s"] for r in rows])
    print(json.dumps(dict(n=n, train=a.train, epochs=a.epochs, label_s=round(tl, 1), train_s=round(tt, 1),
                          loss_first=losses[0], loss_last=losses[-1], vscale=vscale, pscale=pscale,
                          eval=rows, mean_zero_vcycles=zv, mean_lith_vcycles=lv, mean_net_vcycles=nv, mean_zero_s=zs, mean_net_s=ns,
                          vcycle_gain=zv / nv if nv else None, time_gain=zs / ns if ns else None)))


if __name__ == "__main__":
    main()
