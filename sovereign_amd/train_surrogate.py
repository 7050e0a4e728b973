"""Trains the branching-score surrogate on strong-branching data the solver
generates itself, on the GPU, and writes sovereign_amd/data/surrogate_mkp.npz.
Inputs are the candidate features standardised within each node, targets
sqrt(score / best score of the node).

  python -m sovereign_amd.train_surrogate [--instances 2000] [--epochs 60]

Training data: root nodes of seeded knapsack instances (n=150, m=50, the
benchmark's family at reduced size), every fractional candidate scored by
batched strong branching. Held-out instances (other seeds) measure how often
the candidate that full strong branching picks is in the network's top 8,
against the most-fractional order."""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import surrogate as S


def log(msg):
    print(f"[train] {msg}", file=sys.stderr, flush=True)


def node_recall(pred, t, ids, k=8):
    """Nodes (of >= 10 candidates) whose best-scoring candidate is in the top k of pred."""
    hits = 0
    for u in np.unique(ids):
        m = np.nonzero(ids == u)[0]
        if len(m) >= 10:
            hits += m[np.argmax(t[m])] in m[np.argsort(-pred[m], kind="stable")[:k]]
    return int(hits)


def recall_top8(net, seeds):
    hits_s = hits_f = nodes = 0
    for seed in seeds:
        X, t, _ = S.generate_dataset([seed])
        if len(t) < 10:
            continue
        nodes += 1
        best = int(np.argmax(t))
        ys = net.forward(S.node_inputs(X))
        hits_s += best in np.argsort(-ys, kind="stable")[:8]
        hits_f += best in np.argsort(np.abs(X[:, 0] - 0.5), kind="stable")[:8]
    return hits_s, hits_f, nodes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=2000)
    ap.add_argument("--eval", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--out", default=S.WEIGHTS)
    ap.add_argument("--dump", default="", help="also write the dataset (X, t, node ids) to this .npz")
    args = ap.parse_args()
    train_seeds = list(range(100000, 100000 + args.instances))
    eval_seeds = list(range(900000, 900000 + args.eval))
    t0 = time.perf_counter()
    X, t, ids = S.generate_dataset(train_seeds, log=log)
    log(f"dataset: {len(t)} candidates from {len(np.unique(ids))} nodes in {time.perf_counter() - t0:.1f}s")
    if args.dump:
        np.savez(args.dump, X=X, t=t, ids=ids)
        Xe, te, ide = S.generate_dataset(eval_seeds)
        np.savez(args.dump.replace(".npz", "_eval.npz"), X=Xe, t=te, ids=ide)
    # early stopping: 10% of the training nodes validate, in chunks of 2 epochs
    nodes_u = np.unique(ids)
    rng = np.random.default_rng(11)
    val_nodes = rng.choice(nodes_u, size=max(1, len(nodes_u) // 10), replace=False)
    val = np.isin(ids, val_nodes)
    Z = S.dataset_inputs(X, ids)
    net = S.Mlp(S.SHAPE, seed=2024)
    t1 = time.perf_counter()
    best = (-1, None, 0)
    hist = []
    for chunk in range(max(1, args.epochs // 2)):
        hist.extend(net.train(Z[~val], t[~val], epochs=2, batch=256, lr=1e-3, seed=7 + chunk))
        r = node_recall(net.forward(Z[val]), t[val], ids[val])
        log(f"epochs {2 * (chunk + 1)}: train loss {hist[-1]:.4f}, validation top-8 recall {r}")
        if r > best[0]:
            best = (r, net.params, 2 * (chunk + 1))
    net.params = best[1]
    hist = np.array(hist)
    log(f"trained in {time.perf_counter() - t1:.1f}s; kept epoch {best[2]} (validation recall {best[0]})")
    hs, hf, nodes = recall_top8(net, eval_seeds)
    log(f"held-out top-8 recall of the strong-branching choice: surrogate {hs}/{nodes}, most fractional {hf}/{nodes}")
    net.save(args.out, loss=hist, kept_epoch=best[2], train_seeds=np.array(train_seeds),
             eval_seeds=np.array(eval_seeds), eval_recall=np.array([hs, hf, nodes]))
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
