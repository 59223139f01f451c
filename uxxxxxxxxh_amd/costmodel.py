"""Learned cost model for the sweep's tiling: dataset, training, evaluation.

    python -m uxxxxxxxxh_amd.costmodel [--db tune_db.csv] [--out report.json]
        [--problems-per-shape 8] [--configs-per-problem 200] [--epochs 300]

1. collect: for each of the 12 benchmark stencils (FP64), random extents
   (3D: every axis in [64, 1024], at most 2^28 points; 2D: every axis in
   [64, 16384], at most 2^27 points), a random sample of the legal tiling
   space is timed by the C++ tuner (uhs_tune_sweep_list), which appends every
   timing to the tuning database (resumable: timings already there are not
   re-run).  The held-out problems are never part of the training set.
2. train: the 20 -> 3x128 -> 1 MLP (uhs_mlp_*) learns log(ms per sweep) from
   those rows with Adam, on the GPU.
3. evaluate, on the held-out problems (3d3r_box FP64 384x512x640 and
   2d3r_star FP64 12288x6144): exhaustive timed search (the judge) against
   the model's top-k timed, and against the two ways of choosing without the
   network (the hand-set default configuration; k random configurations
   timed).  Reports the slowdown of each pick against the exhaustive best and
   the tuning-time ratio.

The timing, the database, the network and its training are C++/HIP in
libuhs.so; this module only orchestrates them.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import (DT_F64, MLP, Config, DeviceArray, builtin, builtin_names, check, default_config,
               fill_uniform, grid_shape, lib, make_coeffs, mlp_features, tune_list, tune_space,
               tune_sweep, tune_time)

HELD_OUT = [("3d3r_box", [384, 512, 640]), ("2d3r_star", [12288, 6144])]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def random_extents(rng, dims):
    while True:
        if dims == 3:
            e = [int(rng.integers(64, 1025)) for _ in range(3)]
            if np.prod(e) <= 2 ** 28:
                return e
        else:
            e = [int(rng.integers(64, 16385)) for _ in range(2)]
            if np.prod(e) <= 2 ** 27:
                return e


def alloc(case, ext, seed):
    st = builtin(case)
    st = st.with_coeffs(make_coeffs(st, seed))
    h = st.radius
    A = DeviceArray(grid_shape(ext, h), np.float64)
    B = DeviceArray(grid_shape(ext, h), np.float64)
    fill_uniform(A, ext, h, seed)
    fill_uniform(B, ext, h, seed + 1)
    return st, h, A, B


def collect(db, problems_per_shape, configs_per_problem, seed):
    rng = np.random.default_rng(seed)
    n_timed = 0
    t0 = time.time()
    for case in builtin_names():
        dims = 2 if case.startswith("2d") else 3
        for p in range(problems_per_shape):
            ext = random_extents(rng, dims)
            if (case, ext) in HELD_OUT:
                continue
            st, h, A, B = alloc(case, ext, 1000 + p)
            space = tune_space(ext, h, st)
            k = min(configs_per_problem, len(space))
            pick = rng.choice(len(space), k, replace=False)
            r = tune_list(A, B, ext, h, st, [space[i] for i in pick], reps=2, db=db)
            n_timed += k
            log("collect %-10s %-22s %4d configs, best %.3f ms (%s) [%.0f s]"
                % (case, "x".join(map(str, ext)), k, r.best_ms, r.best, time.time() - t0))
            A.close()
            B.close()
    return n_timed


def load_rows(db, exclude=HELD_OUT):
    X, y = [], []
    info = {b: builtin(b) for b in builtin_names()}
    for line in open(db):
        f = line.strip().split(",")
        if len(f) != 15 or f[-1] != "0" or f[1] not in info or f[2] != "0":
            continue
        case, dims = f[1], int(f[3])
        ext = [int(v) for v in f[4:4 + dims]]
        if (case, ext) in exclude:
            continue
        st = info[case]
        cfg = tuple(int(v) for v in f[8:13])
        X.append(mlp_features(dims, st.radius, case.endswith("star"), ext, cfg))
        y.append(np.log(float(f[13])))
    return np.array(X, np.float32), np.array(y, np.float32)


def evaluate(model, topk, seed):
    rng = np.random.default_rng(seed)
    out = {}
    for case, ext in HELD_OUT:
        st, h, A, B = alloc(case, ext, 7)
        space = tune_space(ext, h, st)
        ex = tune_sweep(A, B, ext, h, st, reps=3, db=None)
        check(lib().uhs_tune_clear())
        md = tune_sweep(A, B, ext, h, st, reps=3, db=None, model=model, topk=topk)
        check(lib().uhs_tune_clear())
        dflt = default_config(ext, h, st)
        t_dflt = tune_time(A, B, ext, h, st, dflt, reps=3)
        pick = rng.choice(len(space), topk, replace=False)
        t_rand = time.time()
        rnd = tune_list(A, B, ext, h, st, [space[i] for i in pick], reps=3, db=None)
        t_rand = time.time() - t_rand
        check(lib().uhs_tune_clear())
        res = {
            "space": len(space),
            "exhaustive": {"best": ex.best.tuple(), "ms": ex.best_ms, "tuning_s": ex.seconds},
            "model_topk": {"k": topk, "best": md.best.tuple(), "ms": md.best_ms, "tuning_s": md.seconds,
                           "slowdown_vs_exhaustive": md.best_ms / ex.best_ms - 1.0,
                           "tuning_time_ratio": ex.seconds / md.seconds},
            "without_model_default": {"config": dflt.tuple(), "ms": t_dflt,
                                      "slowdown_vs_exhaustive": t_dflt / ex.best_ms - 1.0},
            "without_model_random_k": {"best": rnd.best.tuple(), "ms": rnd.best_ms, "tuning_s": t_rand,
                                       "slowdown_vs_exhaustive": rnd.best_ms / ex.best_ms - 1.0},
        }
        log("%s %s: exhaustive %.3f ms in %.1f s over %d; model top-%d %.3f ms (%+.1f%%) in %.2f s (x%.0f); "
            "default %+.1f%%; random-%d %+.1f%%"
            % (case, ext, ex.best_ms, ex.seconds, len(space), topk, md.best_ms,
               100 * res["model_topk"]["slowdown_vs_exhaustive"], md.seconds,
               res["model_topk"]["tuning_time_ratio"], 100 * res["without_model_default"]["slowdown_vs_exhaustive"],
               topk, 100 * res["without_model_random_k"]["slowdown_vs_exhaustive"]))
        out["%s_%s" % (case, "x".join(map(str, ext)))] = res
        A.close()
        B.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--db", default="tune_db.csv")
    ap.add_argument("--out", default="costmodel_report.json")
    ap.add_argument("--model", default="costmodel.bin")
    ap.add_argument("--problems-per-shape", type=int, default=8)
    ap.add_argument("--configs-per-problem", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--topk", type=int, default=5)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--skip-collect", action="store_true")
    a = ap.parse_args(argv)
    lib()
    t0 = time.time()
    if not a.skip_collect:
        collect(a.db, a.problems_per_shape, a.configs_per_problem, a.seed)
    t_collect = time.time() - t0
    X, y = load_rows(a.db)
    log("training rows: %d" % len(X))
    idx = np.random.default_rng(a.seed).permutation(len(X))
    n_val = max(1, len(X) // 10)
    val, tr = idx[:n_val], idx[n_val:]
    m = MLP(a.seed)
    t1 = time.time()
    loss = m.train(X[tr], y[tr], epochs=a.epochs, batch=256, lr=1e-3, fit_norm=True, seed=a.seed)
    t_train = time.time() - t1
    pv = m.predict(X[val])
    val_rmse = float(np.sqrt(np.mean((pv - y[val]) ** 2)))
    # Spearman rank correlation on the validation rows
    rk = lambda v: np.argsort(np.argsort(v))
    rho = float(np.corrcoef(rk(pv), rk(y[val]))[0, 1])
    log("train loss %.4f, validation RMSE(log ms) %.3f, rank corr %.3f, train %.1f s" % (loss, val_rmse, rho, t_train))
    m.save(a.model)
    ev = evaluate(m, a.topk, a.seed)
    rep = {"rows": int(len(X)), "train_rows": int(len(tr)), "val_rows": int(n_val), "collect_s": t_collect,
           "train_s": t_train, "epochs": a.epochs, "final_train_loss": loss, "val_rmse_log_ms": val_rmse,
           "val_rank_corr": rho, "held_out": ev}
    with open(a.out, "w") as fh:
        json.dump(rep, fh, indent=1)
    print(json.dumps(rep))
    m.close()


if __name__ == "__main__":
    main()
