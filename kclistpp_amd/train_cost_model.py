"""Train the learned root-cost model on the GPU.

Samples come from timed passes of the plain enumeration kernel
(``Sim.cost_samples``: one row per scheduled root, eight features, labels
log2(1 + build cycles) and log2(1 + counting cycles)) on seeded graphs of
the two Sim-kClist++ workload shapes - Erdos-Renyi n=20000, m=100000 with a
planted 30-clique and Chung-Lu n=1e6, m=1.6e7 (gamma 2.3, w_max 2e4) with a
150-vertex community - at several k.  Training runs the library's own BF16
MFMA forward/backward kernels with Adam (``CostModel.train``).

    python -m kclistpp_amd.train_cost_model --out kclistpp_amd/cost_model.kclcm

``online(sim, model)`` is the same step on a graph's own timings: it
fine-tunes a copy of ``model`` on one timed pass of ``sim`` and replans the
sim with it.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import SCHED_COUNTED, SCHED_LEARNED, CostModel, Graph, Sim
from . import graphgen as gg

PKG = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(PKG, "cost_model.kclcm")

# (workload, seed, scale_down, k values); seeds differ from the benchmark's.
TRAINING_GRAPHS = [
    ("er20k", 11, 1, (3, 4, 5)),
    ("cl1m", 12, 1, (4, 5, 6)),
    ("cl1m", 13, 4, (3, 5, 7)),
]


def _log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def samples_for(graph: Graph, k: int, passes: int = 2) -> tuple[np.ndarray, np.ndarray]:
    """Features and labels of every scheduled root of (graph, k), labels
    averaged (in cycles) over `passes` timed passes of the COUNTED plan."""
    with Sim(graph, k, schedule=SCHED_COUNTED) as s:
        acc = None
        for _ in range(passes):
            _, x, y = s.cost_samples()
            c = np.exp2(y.astype(np.float64)) - 1
            acc = c if acc is None else acc + c
        return x, np.log2(1 + acc / passes).astype(np.float32)


def collect(graphs=TRAINING_GRAPHS, passes: int = 2) -> tuple[np.ndarray, np.ndarray, list]:
    xs, ys, meta = [], [], []
    for name, seed, scale, ks in graphs:
        t0 = time.time()
        pairs = gg.workload(name, seed=seed, scale_down=scale)
        with Graph(pairs) as g:
            _log(f"{name} seed {seed} /{scale}: n={g.n} m={g.m} core={g.core} "
                 f"({time.time() - t0:.1f} s)")
            for k in ks:
                if k > g.core + 1:
                    continue
                t1 = time.time()
                x, y = samples_for(g, k, passes)
                xs.append(x)
                ys.append(y)
                meta.append({"graph": name, "seed": seed, "scale_down": scale, "k": k,
                             "rows": int(len(x))})
                _log(f"  k={k}: {len(x)} roots ({time.time() - t1:.1f} s)")
    return np.concatenate(xs), np.concatenate(ys), meta


def train(x: np.ndarray, y: np.ndarray, seed: int = 1, epochs: int = 60,
          batch: int = 2048, lr: float = 2e-3) -> tuple[CostModel, np.ndarray]:
    m = CostModel(seed=seed)
    losses = m.train(x, y, epochs=epochs, batch=batch, lr=lr, fit_norm=True, seed=seed)
    return m, losses


def online(sim: Sim, model: CostModel | None = None, epochs: int = 20,
           seed: int = 1) -> CostModel:
    """Fine-tune on the sim's own first-pass timings and replan with it."""
    _, x, y = sim.cost_samples()
    m = CostModel(seed=seed)
    if model is not None:
        m.set(*model.get())
    m.train(x, y, epochs=epochs, batch=2048, lr=1e-3 if model is not None else 2e-3,
            fit_norm=model is None, seed=seed)
    sim.replan(SCHED_LEARNED, m)
    return m


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--samples", default=None, help="also write the samples (.npz)")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    x, y, meta = collect()
    _log(f"{len(x)} samples; training")
    t0 = time.time()
    m, losses = train(x, y, seed=a.seed, epochs=a.epochs)
    _log(f"trained in {time.time() - t0:.2f} s; loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    m.save(a.out)
    if a.samples:
        np.savez_compressed(a.samples, x=x, y=y, meta=json.dumps(meta), losses=losses)
    print(json.dumps({"model": a.out, "rows": int(len(x)), "graphs": meta,
                      "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
                      "train_s": time.time() - t0}))
    m.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
