"""The C++ autotuner: exhaustive search, the append-only database (resume),
installation of the winner as the default, and the model-pruned search."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def grids(u, case, ext, seed=1):
    st = u.builtin(case)
    st = st.with_coeffs(u.make_coeffs(st, seed))
    h = st.radius
    A = u.DeviceArray(u.grid_shape(ext, h), np.float64)
    B = u.DeviceArray(u.grid_shape(ext, h), np.float64)
    u.fill_uniform(A, ext, h, seed)
    u.fill_uniform(B, ext, h, seed + 1)
    return st, h, A, B


def test_exhaustive_search_resume_and_install(gpu, tmp_path):
    u = gpu
    u.check(u.lib().uhs_tune_clear())
    ext = [40, 64, 96]
    st, h, A, B = grids(u, "3d1r_star", ext)
    space = u.tune_space(ext, h, st)
    assert len(space) > 100
    db = str(tmp_path / "tune.csv")
    r = u.tune_sweep(A, B, ext, h, st, reps=2, db=db)
    assert r.evaluated + r.failed == len(space) and r.reused == 0 and r.evaluated > 0
    lines = open(db).read().strip().splitlines()
    assert len(lines) == len(space)
    # resume: every timing comes from the database
    r2 = u.tune_sweep(A, B, ext, h, st, reps=2, db=db)
    assert r2.reused == len(space) and r2.best.tuple() == r.best.tuple()
    # the winner is now the default for this problem, also after a reload
    assert u.default_config(ext, h, st).tuple() == r.best.tuple()
    u.check(u.lib().uhs_tune_clear())
    assert u.lib().uhs_tune_load_db(db.encode()) >= 1
    assert u.default_config(ext, h, st).tuple() == r.best.tuple()
    # the tuned configuration is no slower than the heuristic one (timing noise aside)
    u.check(u.lib().uhs_tune_clear())
    t_def = u.tune_time(A, B, ext, h, st, u.default_config(ext, h, st), reps=5)
    t_best = u.tune_time(A, B, ext, h, st, r.best, reps=5)
    assert t_best <= 1.25 * t_def
    A.close(); B.close()


def test_model_pruned_search(gpu, tmp_path):
    u = gpu
    u.check(u.lib().uhs_tune_clear())
    ext = [32, 48, 64]
    st, h, A, B = grids(u, "3d2r_box", ext)
    db = str(tmp_path / "t.csv")
    full = u.tune_sweep(A, B, ext, h, st, reps=1, db=db)
    rows = [l.split(",") for l in open(db).read().strip().splitlines()]
    ok = [r for r in rows if r[-1] == "0"]
    X = np.array([u.mlp_features(3, 2, False, ext, tuple(int(v) for v in r[8:13])) for r in ok])
    y = np.log(np.array([float(r[13]) for r in ok]))
    m = u.MLP(4)
    m.train(X, y, epochs=200, batch=32, lr=2e-3)
    r = u.tune_sweep(A, B, ext, h, st, reps=1, db=None, model=m, topk=5)
    assert r.evaluated == 5
    assert r.best_ms <= 1.5 * full.best_ms
    m.close(); A.close(); B.close()
    u.check(u.lib().uhs_tune_clear())


def _db_rows(path):
    rows = [l.split(",") for l in open(path).read().strip().splitlines()]
    return [r for r in rows if r[-1] == "0" and r[0] != "device"]


def test_model_gains_over_choosing_without_it(gpu, tmp_path):
    """Train the network on the autotuner's timings of three 3d2r_box
    problems, then on a held-out extent: its top-5 (timed) must beat the
    median of 25 random 5-configuration picks (the method without the
    network), and its predicted ranking of the whole space must follow the
    measured one (rank correlation > 0.6)."""
    u = gpu
    u.check(u.lib().uhs_tune_clear())
    rng = np.random.default_rng(5)
    X, y = [], []
    for i, ext in enumerate(([96, 128, 160], [160, 96, 128], [128, 160, 96])):
        st, h, A, B = grids(u, "3d2r_box", ext, seed=20 + i)
        db = str(tmp_path / ("train%d.csv" % i))
        u.tune_sweep(A, B, ext, h, st, reps=3, db=db)
        for r in _db_rows(db):
            X.append(u.mlp_features(3, 2, False, ext, tuple(int(v) for v in r[8:13])))
            y.append(np.log(float(r[13])))
        A.close(); B.close()
    m = u.MLP(6)
    m.train(np.array(X), np.array(y), epochs=200, batch=64, lr=2e-3)

    ext = [128, 128, 128]
    st, h, A, B = grids(u, "3d2r_box", ext, seed=30)
    db = str(tmp_path / "held.csv")
    full = u.tune_sweep(A, B, ext, h, st, reps=3, db=db)
    held = _db_rows(db)
    cfgs = [tuple(int(v) for v in r[8:13]) for r in held]
    ms = np.array([float(r[13]) for r in held])
    pred = m.predict(np.array([u.mlp_features(3, 2, False, ext, c) for c in cfgs]))
    rank = lambda a: np.argsort(np.argsort(a))
    corr = float(np.corrcoef(rank(pred), rank(ms))[0, 1])
    # Both picks are scored on the same exhaustive timings (no re-timing noise).
    model_best = float(np.min(ms[np.argsort(pred)[:5]]))
    random_best = sorted(float(np.min(ms[rng.choice(len(ms), 5, replace=False)])) for _ in range(25))
    median_random = random_best[len(random_best) // 2]
    pruned = u.tune_sweep(A, B, ext, h, st, reps=3, db=None, model=m, topk=5)  # the C++ pruned search
    print("held-out 3d2r_box 128^3 (%d configurations): exhaustive %.4f ms, model top-5 %.4f ms "
          "(pruned search re-timed %.4f ms), median random-5 %.4f ms, rank corr %.3f"
          % (len(ms), full.best_ms, model_best, pruned.best_ms, median_random, corr))
    assert pruned.evaluated == 5
    assert corr > 0.6
    assert model_best <= median_random
    m.close(); A.close(); B.close()
    u.check(u.lib().uhs_tune_clear())
