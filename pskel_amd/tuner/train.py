"""Train the launch cost model on timings measured on this GPU.

    python -m pskel_amd.tuner.train [--train N] [--test M] [--out pskel_amd/data/costmodel.bin]
                                    [--report path.json] [--ensemble E]
                                    [--save-rows rows.json | --load-rows rows.json]

Collects exhaustive timings of every candidate for N random queries (about 20-40
candidates each), trains E MLPs from different initialisations with the
library's HIP kernels and ships their mean (packed into one model file), then
measures M held-out queries exhaustively and reports, per query, the model's
regret (time of its choice / best time - 1), the heuristic's and the GA's
regret and measurement count, and the speed-up of the model's choice over the
plain kernel.  --save-rows keeps both timing tables (gzip when the name ends
in .gz); --load-rows trains and reports from a saved pair instead of measuring
again, timing only candidates added since.  The shipped model is trained from
pskel_amd/data/tuner_rows.json.gz.
"""
from __future__ import annotations

import argparse
import gzip
import json
import math
import sys
import time

import numpy as np

from . import (CostModel, Measurer, candidates, cfg_dict, cfg_key, collect, dataset, exhaustive, genetic,
               pack_ensemble, random_queries, regret, rows_from_json, rows_to_json, SHIPPED_MODEL)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=160)
    ap.add_argument("--test", type=int, default=40)
    ap.add_argument("--epochs", type=int, default=600)
    ap.add_argument("--out", default=SHIPPED_MODEL)
    ap.add_argument("--report", default="")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--ensemble", type=int, default=5, help="members averaged into the shipped model")
    ap.add_argument("--save-rows", default="")
    ap.add_argument("--load-rows", default="")
    args = ap.parse_args(argv)

    log = lambda m: print(m, flush=True)  # noqa: E731
    t0 = time.time()
    def opener(path, mode):
        return gzip.open(path, mode + "t") if path.endswith(".gz") else open(path, mode)

    if args.load_rows:
        with opener(args.load_rows, "r") as f:
            saved = json.load(f)
        rows, miss = rows_from_json(saved["train"])
        test_rows, miss_t = rows_from_json(saved["test"])
        # candidates added since the table was measured: time them now
        for table, missing in ((rows, miss), (test_rows, miss_t)):
            todo = list(dict.fromkeys(q for q, _ in missing))
            for n, q in enumerate(todo):
                m = Measurer(q)
                try:
                    keys = [cfg_key(c) for c in candidates(q)]
                    for qq, c in missing:
                        if qq == q:
                            table.append({"q": q, "ci": keys.index(cfg_key(c)), "ms": m(c)})
                finally:
                    m.close()
                if n % 10 == 9 or n + 1 == len(todo):
                    log(f"measured new candidates on {n + 1}/{len(todo)} queries ({time.time() - t0:.0f} s)")
        log(f"loaded {len(rows)} + {len(test_rows)} timings, measured {len(miss) + len(miss_t)} new candidates")
        train_q = list(dict.fromkeys(r["q"] for r in rows))
        test_q = list(dict.fromkeys(r["q"] for r in test_rows))
        if args.save_rows:
            with opener(args.save_rows, "w") as f:
                json.dump({"train": rows_to_json(rows), "test": rows_to_json(test_rows)}, f)
    else:
        train_q = random_queries(args.train, args.seed)
        rows = collect(train_q, log=log if args.train <= 400 else None)
        test_q = random_queries(args.test, args.seed + 1)
        test_rows = collect(test_q)
        if args.save_rows:
            with opener(args.save_rows, "w") as f:
                json.dump({"train": rows_to_json(rows), "test": rows_to_json(test_rows)}, f)
    X, y = dataset(rows)
    log(f"training set: {len(y)} timings from {len(train_q)} queries ({time.time() - t0:.0f} s)")
    members = []
    Xt, yt = dataset(test_rows)
    member_regret = []
    for j in range(max(1, args.ensemble)):
        m = CostModel(seed=args.seed + 7919 * j)
        loss = m.fit(X, y, epochs=args.epochs, seed=args.seed + 104729 * j)
        pred = m.predict(X)
        log(f"member {j}: train loss {loss:.4f} (standardised), rms log-error "
            f"{np.sqrt(np.mean((pred - y) ** 2)):.3f}")
        member_regret.append(float(np.mean([x["regret"] for x in regret(m, test_rows)])))
        members.append(m)
    model = pack_ensemble(members, *CostModel.normalization(np.asarray(X, np.float32), np.asarray(y, np.float32)))
    for m in members:
        m.close()
    pred = model.predict(X)
    log(f"ensemble of {len(members)}: rms log-error {np.sqrt(np.mean((pred - y) ** 2)):.3f}")
    model.save(args.out)

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
        "ensemble_members": len(members),
        "member_regret_means": member_regret,
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
