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
