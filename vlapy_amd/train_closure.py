"""Train the coarse-velocity closure on the method's own fine-grid runs and
evaluate it against them.

    python -m vlapy_amd.train_closure --out vlapy_amd/data/closure_c5.npz

Training runs: --instances draws (vlapy_amd.closure.draw_instances, seed
--seed) at nx = 64, dt = 0.05 for --steps steps on the fine grid (nv = 4096,
FP64); every --every steps the state is averaged onto nv = 256 and paired
with the per-step correction over --span steps.  Evaluation: --eval-instances
fresh draws (seed --eval-seed) run on the fine grid (FP64), on the coarse
grid (FP32) and on the coarse grid with the closure; the field-energy
history error and the damping-rate error of each coarse run against the fine
one are printed (one JSON line) and stored in the output file.
"""

import argparse
import json
import time

import numpy as np
import torch

from . import closure as CL


def evaluate(net, inst, steps):
    t = CL.DT * (np.arange(steps) + 1)
    out = {}
    t0 = time.perf_counter()
    ref = CL.rollout(inst, CL.NV_FINE, "fp64", steps)
    torch.cuda.synchronize()
    out["fine_fp64_seconds"] = time.perf_counter() - t0
    runs = {"plain": None, "closure": net}
    for name, n in runs.items():
        t0 = time.perf_counter()
        e2 = CL.rollout(inst, CL.NV, "fp32", steps, n)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        drift = CL.rollout.density_drift
        h = CL.history_errors(ref, e2)
        d = CL.damping_errors(inst, ref, e2, t)
        ok = ~np.isnan(d)
        out[name] = {"seconds": sec, "density_drift_max": float(np.max(drift)),
                     "finite": bool(np.all(np.isfinite(e2))),
                     "history_err_median": float(np.median(h)), "history_err_mean": float(np.mean(h)),
                     "damping_fitted": int(ok.sum()), "damping_err_median": float(np.median(d[ok])),
                     "damping_within_1pct": float(np.mean(d[ok] <= 0.01))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--span", type=int, default=20)
    ap.add_argument("--train-steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--eval-instances", type=int, default=128)
    ap.add_argument("--eval-seed", type=int, default=12)
    ap.add_argument("--relative", type=int, default=0, help="1: input relative to the background, per-row scale")
    ap.add_argument("--conserve", type=int, default=1, help="1: corrections add no density, momentum, energy")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    t0 = time.perf_counter()
    x, t = CL.training_pairs(CL.draw_instances(a.instances, a.seed), a.steps, a.every, span=a.span,
                             log=lambda d: print("  fine step %d" % d, flush=True) if d % 500 == 0 else None)
    torch.cuda.synchronize()
    print("pairs %d rows, %.1f s" % (x.shape[0], time.perf_counter() - t0), flush=True)
    t0 = time.perf_counter()
    net, hist = CL.train(x, t, a.train_steps, batch=a.batch, lr=a.lr, seed=a.seed,
                         bg=CL.background() if a.relative else None, conserve=a.conserve,
                         log=lambda it, l: print("  step %d loss %.4f" % (it, l), flush=True))
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del x, t
    res = evaluate(net, CL.draw_instances(a.eval_instances, a.eval_seed), a.steps)
    res.update({"train_seconds": train_s, "final_loss": hist[-1][1], "config": vars(a)})
    net.save(a.out, metrics=json.dumps(res))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
