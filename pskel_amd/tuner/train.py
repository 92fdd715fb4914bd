"""Train the launch cost model on timings measured on this GPU.

    python -m pskel_amd.tuner.train [--train N] [--test M] [--out pskel_amd/data/costmodel.bin]
                                    [--report path.json]

Collects exhaustive timings of every candidate for N random queries (about 20-40
candidates each), trains the MLP with the library's HIP kernels, then measures M
held-out queries exhaustively and reports, per query, the model's regret
(time of its choice / best time - 1), the GA's regret and measurement count,
and the speed-up of the model's choice over the plain kernel.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import (CostModel, Measurer, candidates, cfg_dict, cfg_key, collect, dataset, exhaustive, genetic,
               random_queries, regret, DEFAULT_MODEL)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=160)
    ap.add_argument("--test", type=int, default=40)
    ap.add_argument("--epochs", type=int, default=600)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--report", default="")
    ap.add_argument("--seed", type=int, default=2024)
    args = ap.parse_args(argv)

    log = lambda m: print(m, flush=True)  # noqa: E731
    t0 = time.time()
    train_q = random_queries(args.train, args.seed)
    rows = collect(train_q, log=log if args.train <= 400 else None)
    X, y = dataset(rows)
    log(f"training set: {len(y)} timings from {len(train_q)} queries ({time.time() - t0:.0f} s)")
    model = CostModel(seed=args.seed)
    loss = model.fit(X, y, epochs=args.epochs)
    pred = model.predict(X)
    log(f"train loss {loss:.4f} (standardised), rms log-error {np.sqrt(np.mean((pred - y) ** 2)):.3f}")
    model.save(args.out)

    test_q = random_queries(args.test, args.seed + 1)
    test_rows = collect(test_q)
    Xt, yt = dataset(test_rows)
    pt = model.predict(Xt)
    reg = regret(model, test_rows)
    # GA on the held-out queries (measured objective, cached timings)
    ga = []
    for i, q in enumerate(test_q):
        m = Measurer(q)
        try:
            best_cfg, best_t, all_t = exhaustive(q, m)
            m.cache.clear()
            cfg, t, nmeas = genetic(q, m, seed=args.seed + i)
            ga.append({"regret": t / best_t - 1.0, "measured": nmeas, "candidates": len(all_t)})
        finally:
            m.close()
    r = np.array([x["regret"] for x in reg])
    hr = np.array([x["heuristic_regret"] for x in reg])
    gain = np.array([x["plain_ms"] / x["model_ms"] for x in reg])
    report = {
        "train_queries": len(train_q),
        "train_timings": int(len(y)),
        "test_queries": len(test_q),
        "test_timings": int(len(yt)),
        "test_rms_log_error": float(np.sqrt(np.mean((pt - yt) ** 2))),
        "model_regret_mean": float(r.mean()),
        "model_regret_median": float(np.median(r)),
        "model_regret_p90": float(np.percentile(r, 90)),
        "model_regret_max": float(r.max()),
        "heuristic_regret_mean": float(np.nanmean(hr)),
        "ga_regret_mean": float(np.mean([g["regret"] for g in ga])),
        "ga_measured_mean": float(np.mean([g["measured"] for g in ga])),
        "candidates_mean": float(np.mean([g["candidates"] for g in ga])),
        "model_speedup_over_plain_geomean": float(np.exp(np.mean(np.log(gain)))),
        "per_query": [{"shape": list(x["q"].shape), "stencil": x["q"].stencil_name, "storage": x["q"].storage,
                       "iterations": x["q"].iterations, "best_ms": x["best_ms"], "model_ms": x["model_ms"],
                       "heuristic_ms": x["heuristic_ms"], "plain_ms": x["plain_ms"], "regret": x["regret"],
                       "heuristic_regret": x["heuristic_regret"]} for x in reg],
        "seconds": time.time() - t0,
    }
    log(json.dumps({k: v for k, v in report.items() if k != "per_query"}))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    model.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
