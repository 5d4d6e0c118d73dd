"""Train the learned SCC start-charge model on the GPU and evaluate it.

    python -m gfn1_amd.train_guess [--train 20000] [--eval 4096] [--schedule 40:1e-3] [--out PATH]

Data: noisy snapshots of the mixed H2O/CH4/NH3 clusters (geometry.noisy_snapshot,
seeds 10^6 + j, N(0, 0.05 A) displacements), labelled with the converged shell
charges of the plain FP64 SCC (batch engine).  Training: the library's own HIP
kernels, FP32 Adam, 1024 atoms per step, in stages of (epochs, learning rate)
(--schedule "100:1e-3,100:3e-4"; each stage restarts the Adam moments).  Evaluation: fresh snapshots (indices
after the training set) run with the standard and with the learned start at
the benchmark's SCC tolerance; reports iterations, wall time and the largest
energy difference.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import geometry
from .api import AnalyticGradientOptions, BatchEngine, ElectronicOptions, GuessModel, basis_counts

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data", "guess_mlp.bin")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def label(systems, tol):
    eng = BatchEngine(systems, options=AnalyticGradientOptions(electronic=ElectronicOptions(**tol))).run()
    keep, q = [], []
    for i, (z, _) in enumerate(systems):
        if eng.status(i) == 0:
            keep.append(i)
            q.append(eng.shell_charges(i, basis_counts(z)[0]))
    return [systems[i] for i in keep], q


def evaluate(systems, model, tol):
    opts = AnalyticGradientOptions(electronic=ElectronicOptions(**tol))
    out = {}
    for name, guess in (("standard", None), ("learned", model)):
        eng = BatchEngine(systems, options=opts, guess=guess)
        eng.run()  # warm-up
        t = time.perf_counter()
        eng.run()
        dt = time.perf_counter() - t
        out[name] = {"seconds": dt, "molecules_per_s": len(systems) / dt,
                     "mean_iterations": float(np.mean([eng.iterations(i) for i in range(len(systems))])),
                     "energies": np.array([eng.energy(i) for i in range(len(systems))]),
                     "converged": int(sum(eng.status(i) == 0 for i in range(len(systems))))}
    de = float(np.max(np.abs(out["standard"].pop("energies") - out["learned"].pop("energies"))))
    return out, de


def protocol(n_train=20000, n_eval=4096, schedule="40:1e-3", batch_atoms=1024, seed=1, out=None, n_tight=512):
    """The learned-start protocol: label n_train noisy snapshots with the plain
    SCC, train a fresh model on the GPU, evaluate it on n_eval fresh snapshots
    at the benchmark tolerance (and the first n_tight at the labelling
    tolerance).  Saves the model to `out` when given; returns the report."""
    label_tol = dict(energy_tolerance=1e-10, charge_tolerance=1e-7)
    t0 = time.time()
    train = geometry.noisy_snapshots(n_train)
    log("generated %d training snapshots in %.1fs" % (len(train), time.time() - t0))
    t0 = time.time()
    train, q = label(train, label_tol)
    t_label = time.time() - t0
    log("labelled %d snapshots with the plain SCC in %.1fs" % (len(train), t_label))
    model = GuessModel(seed=seed)
    stages = [(int(a), float(b)) for a, b in (x.split(":") for x in schedule.split(","))]
    t0 = time.time()
    loss = []
    for k, (ep, lr) in enumerate(stages):
        loss += list(model.train(train, q, epochs=ep, batch_atoms=batch_atoms, learning_rate=lr, seed=seed + k))
        log("stage %d: %d epochs at %.1e, loss %.3e" % (k, ep, lr, loss[-1]))
    t_train = time.time() - t0
    epochs = sum(ep for ep, _ in stages)
    log("trained %d epochs in %.1fs, loss %.3e -> %.3e" % (epochs, t_train, loss[0], loss[-1]))
    if out:
        model.save(out)
    t0 = time.time()
    fresh = geometry.noisy_snapshots(n_eval, first=n_train)
    log("generated %d evaluation snapshots in %.1fs" % (len(fresh), time.time() - t0))
    ev, de = evaluate(fresh, model, dict(energy_tolerance=1e-7))
    ev_tight, de_tight = evaluate(fresh[: min(n_tight, len(fresh))], model, label_tol)
    return {
        "train_snapshots": len(train), "train_atoms": int(sum(len(z) for z, _ in train)), "epochs": epochs,
        "batch_atoms": batch_atoms, "schedule": schedule, "label_seconds": t_label, "train_seconds": t_train,
        "loss_first_epoch": float(loss[0]), "loss_last_epoch": float(loss[-1]),
        "eval_snapshots": len(fresh), "eval_bench_tolerance": ev, "max_abs_dE_bench_tolerance": de,
        "eval_tight_tolerance": ev_tight, "max_abs_dE_tight_tolerance": de_tight,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--eval", type=int, default=4096)
    ap.add_argument("--schedule", default="40:1e-3", help="comma-separated epochs:learning_rate stages")
    ap.add_argument("--batch-atoms", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args()
    rep = protocol(args.train, args.eval, args.schedule, args.batch_atoms, args.seed, out=args.out)
    rep["model"] = os.path.abspath(args.out)
    print(json.dumps(rep))


if __name__ == "__main__":
    main()
