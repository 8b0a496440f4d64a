"""Training pipeline of the learned Davidson preconditioner (GPU).

1. Seeded small instances (norb 16-24, the workload generator's rules) are
   solved with the thick-restart solver while it records (residual, Ritz
   vector, Ritz value) per step.
2. Each recorded step gives (features, target) pairs: the determinant's
   features (occupations, residual, h_II - E, diagonal step, excitation
   levels) and the matching component of the unit error vector of that
   step's Ritz vector relative to the converged vector (sbdx_mlp_collect).
3. The MLP is trained on the pooled pairs with the library's own BF16-MFMA
   forward/backward kernels and Adam (sbdx_mlp_train).
The weights are plain float32 arrays (save/load with numpy).
"""
from __future__ import annotations

import time

import numpy as np

from . import TPBOperator, basis_init_vector, instances, lib

TRAIN_INSTANCES = [
    dict(norb=16, nelec_half=5, nstr=400, nchol=40, seed=101),
    dict(norb=18, nelec_half=5, nstr=500, nchol=40, seed=102),
    dict(norb=20, nelec_half=6, nstr=500, nchol=60, seed=103),
    dict(norb=24, nelec_half=6, nstr=600, nchol=80, seed=104),
]


def init_weights(seed=0):
    """He-normal hidden layers, small output layer, zero biases."""
    n = lib().sbdx_mlp_nparam()
    din, hid = lib().sbdx_mlp_nfeat(), 128
    rng = np.random.default_rng(seed)
    p = np.zeros(n, dtype=np.float32)
    o = 0
    p[o:o + din * hid] = rng.normal(0, np.sqrt(2.0 / din), din * hid)
    o += din * hid + hid
    p[o:o + hid * hid] = rng.normal(0, np.sqrt(2.0 / hid), hid * hid)
    o += hid * hid + hid
    p[o:o + hid] = rng.normal(0, np.sqrt(1.0 / hid), hid)
    return p


def generate_pairs(specs=None, pairs_per_instance=500_000, record=256, seed=0, log=print):
    feats, tgts = [], []
    for k, spec in enumerate(specs or TRAIN_INSTANCES):
        inst = instances.make_instance(**spec)
        op = TPBOperator.from_instance(inst)
        op.configure("thick")
        op.mlp_record(record)
        W0 = np.random.default_rng(spec["seed"]).standard_normal(op.ndet)
        W0 /= np.linalg.norm(W0)
        op.set_vector(W0)
        t0 = time.perf_counter()
        E, n, ok = op.solve(max_it=200, nb=12, eps=1e-8)
        f, t = op.mlp_collect(pairs_per_instance, seed=seed + k)
        log(f"[learned] instance {spec}: ndet={op.ndet} E={E:.10f} steps={n} converged={ok} "
            f"pairs={t.size} ({time.perf_counter() - t0:.1f} s)")
        op.close()
        feats.append(f)
        tgts.append(t)
    return np.concatenate(feats), np.concatenate(tgts)


def train(op: TPBOperator, feat, tgt, epochs=20, batch=4096, lr=1e-3, seed=0, weights=None):
    op.mlp_set_weights(init_weights(seed) if weights is None else weights)
    loss = op.mlp_train(feat, tgt, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return op.mlp_get_weights(), loss


def baseline_loss(feat, tgt):
    """MSE of the best scalar multiple of the diagonal step (feature 130)
    -- what the network has to beat to add anything over R/(E0-d)."""
    d = feat[:, 130].astype(np.uint32) << 16
    x = d.view(np.float32).astype(np.float64)
    a = float(x @ tgt / max(x @ x, 1e-30))
    return float(np.mean((a * x - tgt) ** 2)), float(np.mean(tgt.astype(np.float64) ** 2))


EVAL_INSTANCE = dict(norb=22, nelec_half=6, nstr=700, nchol=60, seed=211)  # held out


def evaluate(weights, spec=None, inst=None, solver="thick", eps=1e-6, max_it=400, nb=12,
             start="hf", log=print):
    """Time-to-|r|<eps and sigma products of the thick-restart solver with the
    diagonal preconditioner and with the MLP (weights), same start vector.
    Returns {"diag": {...}, "mlp": {...}, "gain_products": diag/mlp,
    "gain_time": diag/mlp}."""
    if inst is None:
        inst = instances.make_instance(**(spec or EVAL_INSTANCE))
    op = TPBOperator.from_instance(inst)
    if start == "hf":
        W0 = basis_init_vector(op.ndet)
    else:
        W0 = np.random.default_rng(7).standard_normal(op.ndet)
        W0 /= np.linalg.norm(W0)
    op.mlp_set_weights(weights)
    out = {}
    for mode, name in ((0, "diag"), (1, "mlp")):
        op.configure(solver)
        op.set_preconditioner(mode)
        op.set_vector(W0)
        t0 = time.perf_counter()
        E, n, ok = op.solve(max_it=max_it, nb=nb, eps=eps)
        dt = time.perf_counter() - t0
        out[name] = dict(E=E, products=n, converged=ok, time_s=round(dt, 4),
                         steps_by_precond=op.precond_stats())
        log(f"[learned] eval {name}: E={E:.12f} products={n} converged={ok} time={dt:.3f} s "
            f"{out[name]['steps_by_precond']}")
    op.set_preconditioner(0)
    op.close()
    out["dE"] = abs(out["mlp"]["E"] - out["diag"]["E"])
    out["gain_products"] = out["diag"]["products"] / max(out["mlp"]["products"], 1)
    out["gain_time"] = out["diag"]["time_s"] / max(out["mlp"]["time_s"], 1e-9)
    return out


def _main(argv=None):
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="train / evaluate the learned preconditioner")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True, help="weights file (.npy, float32)")
    t.add_argument("--epochs", type=int, default=20)
    t.add_argument("--pairs", type=int, default=500_000, help="pairs per training instance")
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--seed", type=int, default=0)
    e = sub.add_parser("eval")
    e.add_argument("--weights", required=True)
    e.add_argument("--eps", type=float, default=1e-6)
    e.add_argument("--solver", default="thick")
    args = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    if args.cmd == "train":
        t0 = time.perf_counter()
        feat, tgt = generate_pairs(pairs_per_instance=args.pairs, seed=args.seed, log=log)
        base, var = baseline_loss(feat, tgt)
        inst = instances.make_instance(**TRAIN_INSTANCES[0])
        op = TPBOperator.from_instance(inst)
        w, loss = train(op, feat, tgt, epochs=args.epochs, lr=args.lr, seed=args.seed)
        op.close()
        np.save(args.out, w)
        res = dict(pairs=int(tgt.size), loss=[float(x) for x in loss], target_var=var,
                   diag_step_loss=base, seconds=round(time.perf_counter() - t0, 2))
        log(f"[learned] trained on {tgt.size} pairs: loss {loss[0]:.4g} -> {loss[-1]:.4g} "
            f"(diagonal step {base:.4g}, target variance {var:.4g})")
        print(json.dumps(res))
    else:
        w = np.load(args.weights)
        print(json.dumps(evaluate(w, solver=args.solver, eps=args.eps, log=log)))


if __name__ == "__main__":
    _main()
