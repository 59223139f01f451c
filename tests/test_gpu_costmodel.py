"""The learned cost model on its own workload: the 20-input MLP trained on
the GPU by the library's kernels on timings the C++ autotuner collects
(random problems of all 12 stencils, FP64 and both FP32 forms), judged on the
two held-out problems -- 3d3r_box FP64 384x512x640 and 2d3r_star FP64
12288x6144 -- never seen in training.  Three models trained from three
seeds rank the space by the mean of their predicted log times (one model's
ranking among the 2D K-step depths, whose best tiles lie within 5-10 % of
each other, swings with the seed and the timing noise of the collection).
The ensemble's top-5, timed, must come
within 10 % of the exhaustive timed search's best with at least 20x less
tuning time; its ranking of the whole space must follow the measured one.

The batched graph's model: trained on random other batches, its top-5 on a
held-out batch must come within 10 % of that batch's exhaustive best."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.timeout(900)
def test_costmodel_held_out_pair(gpu, tmp_path):
    u = gpu
    from uxxxxxxxxh_amd import costmodel as cm
    db = str(tmp_path / "tune_db.csv")
    n = cm.collect(db, problems_per_shape=6, configs_per_problem=150, seed=11)
    X, y = cm.load_rows(db)
    log("cost model: %d timed configurations, %d rows" % (n, len(X)))
    assert len(X) > 5000 and set(np.unique(X[:, 17]).tolist()) == {0.0, 1.0, 2.0}
    models = [u.MLP(s) for s in (11, 12, 13)]
    try:
        for s, m in zip((11, 12, 13), models):
            m.train(X, y, epochs=300, batch=256, lr=1e-3, fit_norm=True, seed=s)
        ev = cm.evaluate(models, topk=5, seed=11)
    finally:
        for m in models:
            m.close()
        u.check(u.lib().uhs_tune_clear())
    assert set(ev) == {"3d3r_box_384x512x640", "2d3r_star_12288x6144"}
    for name, r in ev.items():
        log(name, r)
        # The picks re-timed under one protocol (the search's own best is the
        # minimum of many noisy timings and reads low), and the same timings
        # of the exhaustive table for both.
        assert r["retimed"]["model_topk_slowdown"] <= 0.10, name
        assert r["in_exhaustive_table"]["model_topk_slowdown"] <= 0.10, name
        assert r["model_topk"]["tuning_time_ratio"] >= 20.0, name


@pytest.mark.timeout(900)
def test_batch_costmodel_held_out_batch(gpu):
    u = gpu
    from uxxxxxxxxh_amd import costmodel as cm
    X, y, _ = cm.batch_collect(n_batches=6, configs_per_batch=60, probe_T=4, seed=5)
    log("batch cost model: %d rows" % len(X))
    assert len(X) >= 300
    m, loss, _ = cm.batch_train(X, y)
    # A held-out batch: seeds and size unlike any training batch.
    specs = cm.random_batch(np.random.default_rng(999), 77)
    grids = cm.batch_grids(specs)
    try:
        ev = cm.batch_evaluate(m, grids, probe_T=4)
    finally:
        m.close()
        cm.close_grids(grids)
    ev.pop("_model_best"), ev.pop("_exhaustive_best")
    log("held-out batch:", ev)
    assert ev["model_topk"]["slowdown_in_exhaustive_table"] <= 0.10
    assert ev["model_topk"]["tuning_time_ratio"] >= 20.0
