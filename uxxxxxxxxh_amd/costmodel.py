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

from . import (DT_F32, DT_F32K, DT_F64, MLP, Config, DeviceArray, batch_default_config, batch_features,
               batch_space, batch_tune_list, batch_tune_model, builtin, builtin_names, check, default_config,
               fill_uniform, grid_shape, lib, make_coeffs, mlp_features, tune_list, tune_space,
               tune_sweep, tune_time)
from .workloads import BATCH_EXTENTS

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


def alloc(case, ext, seed, dtype=DT_F64):
    st = builtin(case)
    st = st.with_coeffs(make_coeffs(st, seed))
    h = st.radius
    npdt = np.float64 if dtype == DT_F64 else np.float32
    A = DeviceArray(grid_shape(ext, h), npdt)
    B = DeviceArray(grid_shape(ext, h), npdt)
    fill_uniform(A, ext, h, seed)
    fill_uniform(B, ext, h, seed + 1)
    return st, h, A, B


def collect(db, problems_per_shape, configs_per_problem, seed, f32_problems=1):
    """Timed random samples of the tiling space: problems_per_shape FP64
    problems per stencil, and f32_problems each in plain and compensated
    FP32 (the precision is a model input)."""
    rng = np.random.default_rng(seed)
    n_timed = 0
    t0 = time.time()
    for case in builtin_names():
        dims = 2 if case.startswith("2d") else 3
        jobs = [DT_F64] * problems_per_shape + [DT_F32, DT_F32K] * f32_problems
        for p, dt in enumerate(jobs):
            ext = random_extents(rng, dims)
            if (case, ext) in HELD_OUT:
                continue
            st, h, A, B = alloc(case, ext, 1000 + p, dt)
            space = tune_space(ext, h, st, dt)
            k = min(configs_per_problem, len(space))
            pick = rng.choice(len(space), k, replace=False)
            r = tune_list(A, B, ext, h, st, [space[i] for i in pick], dtype=dt, reps=2, db=db)
            n_timed += k
            log("collect %-10s %-22s dtype %d %4d configs, best %.3f ms (%s) [%.0f s]"
                % (case, "x".join(map(str, ext)), dt, k, r.best_ms, r.best, time.time() - t0))
            A.close()
            B.close()
    return n_timed


def load_rows(db, exclude=HELD_OUT):
    X, y = [], []
    info = {b: builtin(b) for b in builtin_names()}
    for line in open(db):
        f = line.strip().split(",")
        if len(f) != 15 or f[-1] != "0" or f[1] not in info or f[2] not in ("0", "1", "2"):
            continue
        case, dtype, dims = f[1], int(f[2]), int(f[3])
        ext = [int(v) for v in f[4:4 + dims]]
        if dtype == DT_F64 and (case, ext) in exclude:
            continue
        st = info[case]
        cfg = tuple(int(v) for v in f[8:13])
        X.append(mlp_features(dims, st.radius, case.endswith("star"), ext, cfg, dtype=dtype))
        y.append(np.log(float(f[13])))
    return np.array(X, np.float32), np.array(y, np.float32)


def _table(db, space):
    """Every configuration's time from a database written by one search."""
    t = {}
    for line in open(db):
        f = line.strip().split(",")
        if len(f) == 15 and f[-1] == "0":
            t[tuple(int(v) for v in f[8:13])] = float(f[13])
    return np.array([t.get(c.tuple(), np.inf) for c in space])


def retime(A, B, ext, h, st, cfgs, rounds=7, reps=5):
    """Median ms per sweep of each configuration over `rounds` interleaved
    rounds of `reps` sweeps: one protocol for every method's pick, so the
    comparison carries no selection bias (a search's own best is the minimum
    of many noisy timings, and reads low)."""
    t = [[] for _ in cfgs]
    for _ in range(rounds):
        for i, c in enumerate(cfgs):
            t[i].append(tune_time(A, B, ext, h, st, c, reps=reps))
    return [float(np.median(v)) for v in t]


def evaluate(model, topk, seed, draws=200):
    """Per held-out problem: the exhaustive timed search (the yardstick), the
    model's top-k timed (`model` is one model or a list ranked by the mean
    of their predicted log times), the hand-set default and one random k-sample timed;
    read from the exhaustive search's own table (the same timings for every
    method), the model's top-k and the median of `draws` random k-samples;
    and the four picks re-timed under one protocol."""
    import tempfile
    models = list(model) if isinstance(model, (list, tuple)) else [model]
    rng = np.random.default_rng(seed)
    out = {}
    for case, ext in HELD_OUT:
        st, h, A, B = alloc(case, ext, 7)
        space = tune_space(ext, h, st)
        with tempfile.TemporaryDirectory() as d:
            db = os.path.join(d, "exhaustive.csv")
            ex = tune_sweep(A, B, ext, h, st, reps=3, db=db)
            table = _table(db, space)
        check(lib().uhs_tune_clear())
        feats = np.array([mlp_features(len(ext), st.radius, case.endswith("star"), ext, c) for c in space])
        pred = np.mean([m.predict(feats) for m in models], axis=0)
        top = np.argsort(pred)[:topk]
        valid = np.flatnonzero(np.isfinite(table))
        rand = [table[rng.choice(valid, topk, replace=False)].min() for _ in range(draws)]
        tbest = table[valid].min()
        md = tune_sweep(A, B, ext, h, st, reps=3, db=None, model=model, topk=topk)
        check(lib().uhs_tune_clear())
        dflt = default_config(ext, h, st)
        t_dflt = tune_time(A, B, ext, h, st, dflt, reps=3)
        pick = rng.choice(len(space), topk, replace=False)
        t_rand = time.time()
        rnd = tune_list(A, B, ext, h, st, [space[i] for i in pick], reps=3, db=None)
        t_rand = time.time() - t_rand
        check(lib().uhs_tune_clear())
        rt = retime(A, B, ext, h, st, [ex.best, md.best, dflt, rnd.best])
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
            "in_exhaustive_table": {"model_topk_slowdown": float(table[top].min() / tbest - 1.0),
                                    "random_k_median_slowdown": float(np.median(rand) / tbest - 1.0),
                                    "random_k_draws": draws},
            # The four picks timed again under one protocol (7 interleaved
            # rounds of 5 sweeps, median).
            "retimed": {"exhaustive_ms": rt[0], "model_topk_ms": rt[1], "default_ms": rt[2], "random_k_ms": rt[3],
                        "model_topk_slowdown": rt[1] / rt[0] - 1.0, "default_slowdown": rt[2] / rt[0] - 1.0,
                        "random_k_slowdown": rt[3] / rt[0] - 1.0},
        }
        log("%s %s: exhaustive %.3f ms in %.1f s over %d; model top-%d %.3f ms (%+.1f%%) in %.2f s (x%.0f); "
            "default %+.1f%%; random-%d %+.1f%%; in the exhaustive table: model %+.1f%%, random median %+.1f%%; "
            "re-timed: model %+.1f%%, default %+.1f%%, random %+.1f%%"
            % (case, ext, ex.best_ms, ex.seconds, len(space), topk, md.best_ms,
               100 * res["model_topk"]["slowdown_vs_exhaustive"], md.seconds,
               res["model_topk"]["tuning_time_ratio"], 100 * res["without_model_default"]["slowdown_vs_exhaustive"],
               topk, 100 * res["without_model_random_k"]["slowdown_vs_exhaustive"],
               100 * res["in_exhaustive_table"]["model_topk_slowdown"],
               100 * res["in_exhaustive_table"]["random_k_median_slowdown"],
               100 * res["retimed"]["model_topk_slowdown"], 100 * res["retimed"]["default_slowdown"],
               100 * res["retimed"]["random_k_slowdown"]))
        out["%s_%s" % (case, "x".join(map(str, ext)))] = res
        A.close()
        B.close()
    return out


# ---- the batched graph ----------------------------------------------------

def batch_grids(specs):
    """Device grids (A, B, ext, halo, stencil) for batch specifications
    {case, coeff_seed, ext, seed_a, seed_b} (uxxxxxxxxh_amd.workloads)."""
    grids = []
    for g in specs:
        st = builtin(g["case"])
        st = st.with_coeffs(make_coeffs(st, g["coeff_seed"]))
        h = st.radius
        A = DeviceArray(grid_shape(g["ext"], h), np.float64)
        B = DeviceArray(grid_shape(g["ext"], h), np.float64)
        fill_uniform(A, g["ext"], h, g["seed_a"])
        fill_uniform(B, g["ext"], h, g["seed_b"])
        grids.append((A, B, g["ext"], h, st))
    return grids


def close_grids(grids):
    for A, B, _, _, _ in grids:
        A.close()
        B.close()


def random_batch(rng, k):
    """A training batch: 16-64 grids, a quarter to three quarters of them
    7-point stars and the rest 125-point boxes, extents from the benchmark's
    set, seeds apart from the benchmark batch's."""
    n = int(rng.integers(16, 65))
    nstar = int(round(n * float(rng.choice([0.25, 0.5, 0.75]))))
    specs = []
    for i in range(n):
        specs.append({"case": "3d1r_star" if i < nstar else "3d2r_box", "coeff_seed": 20000 + 100 * k + i,
                      "ext": [int(rng.choice(BATCH_EXTENTS)) for _ in range(3)],
                      "seed_a": 30000 + 100 * k + i, "seed_b": 40000 + 100 * k + i})
    return specs


def batch_collect(n_batches=8, configs_per_batch=80, probe_T=4, seed=5):
    """Times a random sample of the batch-configuration space on random
    batches; rows are (features, log(ms per step))."""
    rng = np.random.default_rng(seed)
    space = batch_space()
    X, y = [], []
    t0 = time.time()
    for k in range(n_batches):
        specs = random_batch(rng, k)
        grids = batch_grids(specs)
        pick = rng.choice(len(space), min(configs_per_batch, len(space)), replace=False)
        cands = [space[i] for i in pick]
        _, bms, ms = batch_tune_list(grids, probe_T, cands)
        for c, t in zip(cands, ms):
            if t > 0:
                X.append(batch_features(grids, c))
                y.append(np.log(t / probe_T))
        log("batch collect %d: %d grids, %d configs, best %.3f ms per step [%.0f s]"
            % (k, len(grids), len(cands), bms / probe_T, time.time() - t0))
        close_grids(grids)
    return np.array(X, np.float32), np.array(y, np.float32), time.time() - t0


def batch_train(X, y, seed=3, epochs=400):
    m = MLP(seed)
    t0 = time.time()
    loss = m.train(X, y, epochs=epochs, batch=64, lr=1e-3, fit_norm=True, seed=seed)
    return m, loss, time.time() - t0


def batch_evaluate(model, grids, probe_T=8, topk=5, seed=7):
    """On a batch the model never saw: the exhaustive timed search (the
    judge) against the model's top-k timed, the hand-set default and k random
    candidates (both read from the exhaustive table, 200 random draws)."""
    space = batch_space()
    t0 = time.time()
    _, ex_ms, ms = batch_tune_list(grids, probe_T, space)
    t_ex = time.time() - t0
    ms = np.array(ms)
    ok = ms > 0
    best_i = int(np.argmin(np.where(ok, ms, np.inf)))
    mb, m_ms, m_sec = batch_tune_model(grids, probe_T, model, topk)
    feats = np.array([batch_features(grids, c) for c in space])
    pred = model.predict(feats)
    top = np.argsort(pred)[:topk]
    rng = np.random.default_rng(seed)
    valid = np.flatnonzero(ok)
    rand = [ms[rng.choice(valid, topk, replace=False)].min() for _ in range(200)]
    d = batch_default_config()
    di = [i for i, c in enumerate(space) if bytes(c) == bytes(d)]
    rk = lambda v: np.argsort(np.argsort(v))
    rho = float(np.corrcoef(rk(pred[ok]), rk(ms[ok]))[0, 1])
    return {
        "space": len(space), "probe_steps": probe_T,
        "exhaustive": {"best": repr(space[best_i]), "ms": float(ms[best_i]), "tuning_s": round(t_ex, 2)},
        "model_topk": {"k": topk, "best": repr(mb), "ms": m_ms, "tuning_s": round(m_sec, 3),
                       "slowdown_vs_exhaustive": m_ms / float(ms[best_i]) - 1.0,
                       "slowdown_in_exhaustive_table": float(ms[top].min() / ms[best_i] - 1.0),
                       "tuning_time_ratio": t_ex / m_sec, "rank_corr": rho},
        "without_model_default": {"slowdown_vs_exhaustive": float(ms[di[0]] / ms[best_i] - 1.0) if di else None},
        "without_model_random_k": {"median_slowdown_vs_exhaustive": float(np.median(rand) / ms[best_i] - 1.0)},
        "_model_best": mb, "_exhaustive_best": space[best_i],
    }


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
    ap.add_argument("--ensemble", type=int, default=3, help="models trained from seeds seed, seed+1, ...")
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
    models = [MLP(a.seed + i) for i in range(max(1, a.ensemble))]
    t1 = time.time()
    losses = [m.train(X[tr], y[tr], epochs=a.epochs, batch=256, lr=1e-3, fit_norm=True, seed=a.seed + i)
              for i, m in enumerate(models)]
    loss = float(np.mean(losses))
    t_train = time.time() - t1
    pv = np.mean([m.predict(X[val]) for m in models], axis=0)
    val_rmse = float(np.sqrt(np.mean((pv - y[val]) ** 2)))
    # Spearman rank correlation on the validation rows
    rk = lambda v: np.argsort(np.argsort(v))
    rho = float(np.corrcoef(rk(pv), rk(y[val]))[0, 1])
    log("train loss %.4f, validation RMSE(log ms) %.3f, rank corr %.3f, train %.1f s" % (loss, val_rmse, rho, t_train))
    for i, m in enumerate(models):
        m.save(a.model if i == 0 else "%s.%d" % (a.model, i))
    ev = evaluate(models, a.topk, a.seed)
    rep = {"rows": int(len(X)), "train_rows": int(len(tr)), "val_rows": int(n_val), "collect_s": t_collect,
           "train_s": t_train, "ensemble": len(models), "epochs": a.epochs, "final_train_loss": loss, "val_rmse_log_ms": val_rmse,
           "val_rank_corr": rho, "held_out": ev}
    with open(a.out, "w") as fh:
        json.dump(rep, fh, indent=1)
    print(json.dumps(rep))
    for m in models:
        m.close()


if __name__ == "__main__":
    main()
