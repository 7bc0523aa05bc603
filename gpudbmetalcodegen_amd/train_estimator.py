"""Train the group-count estimator on the GPU.

1. Log training runs with the engine: device-generated Zipf keys over
   log-uniform key ranges (2^6..2^25), skew 0..1.5, selectivity 1..100%,
   n log-uniform 2^14..2^24 rows; features from an 8,192-row sample; the true
   group count from the engine's exact GB_COUNT plan.
2. Train the MLP with the project's HIP kernels (per-sample backward,
   deterministic reduction + Adam).
3. Save the weights next to this file and the run log (train + held-out) as
   estimator_runs.npz.

    python -m gpudbmetalcodegen_amd.train_estimator [--runs 5000] [--heldout 600] [--epochs 300]
"""

from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

from . import engine as E
from . import estimator as EST

HERE = Path(__file__).resolve().parent
RUNS = HERE / "estimator_runs.npz"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=5000)
    ap.add_argument("--heldout", type=int, default=600)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--out", default=str(EST.WEIGHTS))
    ap.add_argument("--runs-out", default=str(RUNS))
    args = ap.parse_args(argv)
    ctx = E.Context(0)
    t0 = time.time()
    parts = []
    for i, (n, seed) in enumerate([(args.runs, 2024), (args.heldout, 4242)]):
        chunks = []
        done = 0
        while done < n:
            m = min(500, n - done)
            chunks.append(EST.collect_runs(ctx, m, seed * 1000 + done, nmin=1 << 14, nmax=1 << 24))
            done += m
            print(f"[train] logged {done}/{n} runs ({'train' if i == 0 else 'held-out'}) {time.time() - t0:.1f}s",
                  flush=True)
        parts.append({k: np.concatenate([c[k] for c in chunks]) for k in chunks[0]})
    tr, ho = parts
    y = EST.target(tr["groups"], tr["ub"])
    est = EST.CardinalityEstimator(ctx)
    est.init(7)
    t1 = time.time()
    losses = est.train(tr["X"], y, epochs=args.epochs, batch=256, lr=1e-3, seed=11)
    print(f"[train] {args.epochs} epochs in {time.time() - t1:.1f}s; loss {losses[0]:.4f} -> {losses[-1]:.4f}",
          flush=True)
    est.save(args.out)
    pred = est.predict_groups(ho["X"], ho["ub"])
    q = np.maximum(pred / np.maximum(1, ho["groups"]), np.maximum(1, ho["groups"]) / pred)
    print(f"[train] held-out q-error: median {np.median(q):.3f}, p90 {np.percentile(q, 90):.3f}, "
          f"max {q.max():.3f}", flush=True)
    np.savez_compressed(args.runs_out, train_X=tr["X"], train_ub=tr["ub"], train_groups=tr["groups"],
                        train_spec=tr["spec"], heldout_X=ho["X"], heldout_ub=ho["ub"], heldout_groups=ho["groups"],
                        heldout_spec=ho["spec"], losses=losses)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
