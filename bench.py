"""Benchmark of the MI355X stencil sweep (one GPU).

Workload (the headline of the reference's GPU benchmark set): the 3D 7-point
star (3d1r_star) in FP64 on a 512^3 interior with a zero halo of width 1,
coefficients symmetric U[0,1) normalised to sum 1, interior U[0,1), T = 100
ping-pong timesteps.  One "step" is one call of uhs_sweep_f64 over that grid
(100 sweeps); inputs are resident in HBM before the timed region starts.

value = interior points * T * steps / time  (GStencils/s), timed with HIP
events on the stream the sweeps are launched on.  The JSON line also
carries the roofline of the sweep kernel (16 algorithmic bytes per update
against 8 TB/s), the CPU baseline (the oracle's single-thread restatement
of the reference's serial sweep, on a bounded sample), and per-shape
results for all 12 benchmark stencils (512^3 / 8192^2, T = 100).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-shapes]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0        # MI355X HBM3E, published
PEAK_FP64_TFLOPS = 78.6       # MI355X FP64 vector, published


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_sweeps(u, case, ext, T, steps, warmup, seed=1, cfg=None, dtype="f64"):
    """Returns (ms per step, ms per sweep, config) for uhs_sweep_f64 (cfg None:
    the library's default configuration); dtype "f32"/"f32k": uhs_sweep_f32,
    plain or compensated."""
    st = u.builtin(case)
    st = st.with_coeffs(u.make_coeffs(st, seed))
    h = st.radius
    shape = u.grid_shape(ext, h)
    npdt = np.float64 if dtype == "f64" else np.float32
    A = u.DeviceArray(shape, npdt)
    B = u.DeviceArray(shape, npdt)
    u.fill_uniform(A, ext, h, seed)
    u.fill_uniform(B, ext, h, seed + 1)
    run_cfg = cfg
    cdt = {"f64": u.DT_F64, "f32": u.DT_F32, "f32k": u.DT_F32K}[dtype]
    cfg = cfg if cfg is not None else u.default_config(ext, h, st, cdt)

    def sweep():
        if dtype == "f64":
            u.sweep_f64(A, B, ext, h, st, T, run_cfg)
        else:
            u.sweep_f32(A, B, ext, h, st, T, dtype == "f32k", run_cfg)
    for _ in range(warmup):
        sweep()
    u.synchronize()
    tm = u.Timer()
    tm.start()
    for _ in range(steps):
        sweep()
    tm.stop()
    ms = tm.ms()
    u.synchronize()
    A.close()
    B.close()
    return ms / steps, ms / (steps * T), cfg


def one_step(u, n, T):
    """The same workload with one launch per timestep (the one-step kernel
    and its default tile, as the reference runs it): GStencils/s, its
    fraction of 8 TB/s at 16 B per update, and the DRAM bytes per launch
    measured by the traffic counters (committed summary)."""
    ext = [n, n, n]
    st = u.builtin("3d1r_star")
    cfg = u.one_step_config(ext, st.radius, st)
    ms_step, ms_sweep, _ = time_sweeps(u, "3d1r_star", ext, T, 2, 1, cfg=cfg)
    g = float(n) ** 3 / (ms_sweep * 1e-3) / 1e9
    out = {"workload": "3d1r_star FP64 %d^3, T=%d, one step per launch" % (n, T),
           "config": list(cfg.tuple()), "kernel": kernel_name(cfg), "GStencils_s": round(g, 2),
           "avg_launch_ms": round(ms_sweep, 5),
           "algorithmic_GBps_at_16B": round(16.0 * g, 1), "frac_of_8TBps_at_16B": round(16.0 * g / PEAK_HBM_GBPS, 4)}
    tr = load_traffic("3d1r_star", ext)
    if tr:
        out["dram_bytes_per_launch"] = tr["bytes_per_launch"]
        out["dram_GBps"] = round(tr["bytes_per_launch"] / (ms_sweep * 1e-3) / 1e9, 1)
    return out


XGMI_LINK_GBPS = 153.0       # MI355X, one xGMI link, one direction, published


def shard_share(u, T=100, gpus=8):
    """One GPU's share of the 512^3 3d1r_star grid slab-decomposed along
    axis 0 over `gpus` GPUs: the sweep of a 512x512x(512/gpus) slab (its
    ghost plane is the halo), measured, beside the time to exchange one
    2 MiB face per neighbour per step over one xGMI link each (a model: the
    pool has one GPU)."""
    n0 = 512 // gpus
    ms_step, ms_sweep, cfg = time_sweeps(u, "3d1r_star", [n0, 512, 512], T, 2, 1)
    face_bytes = 512 * 512 * 8
    exch_us = face_bytes / (XGMI_LINK_GBPS * 1e9) * 1e6
    return {"workload": "3d1r_star FP64 %dx512x512 slab (one of %d), T=%d" % (n0, gpus, T),
            "config": [cfg.tile_x, cfg.tile_y, cfg.tile_z, cfg.vec, cfg.flags],
            "GStencils_s": round(n0 * 512 * 512 / (ms_sweep * 1e-3) / 1e9, 2),
            "sweep_us": round(ms_sweep * 1e3, 2),
            "face_exchange_us_model": round(exch_us, 2),
            "exchange_over_sweep": round(exch_us / (ms_sweep * 1e3), 3)}


def fma_rate(u, n=256 * 32 * 256, iters=2048, reps=5):
    """Rate of the library's FP64 FMA probe (32 independent chains per
    thread, no memory traffic), in TFLOP/s at 2 flops per FMA: the FP64
    vector rate this device reaches, for the FP64-ALU-bound stencils."""
    out = u.DeviceArray([n], np.float64)
    L = u.lib()
    x, y = 1.0 - 2.0 ** -20, 2.0 ** -21
    for _ in range(2):
        u.check(L.uhs_fma_f64(out.ptr, n, x, y, iters, None))
    u.synchronize()
    tm = u.Timer()
    tm.start()
    for _ in range(reps):
        u.check(L.uhs_fma_f64(out.ptr, n, x, y, iters, None))
    tm.stop()
    ms = tm.ms()
    out.close()
    return 2.0 * 32 * iters * n * reps / (ms * 1e-3) / 1e12


def copy_bandwidth(u, nbytes=2 << 30, reps=10):
    """Rate of the library's streaming copy kernel (read + write bytes / s):
    the HBM bandwidth this device reaches, for context next to the roofline."""
    n = nbytes // 8
    a = u.DeviceArray([n], np.float64)
    b = u.DeviceArray([n], np.float64)
    L = u.lib()
    for _ in range(2):
        u.check(L.uhs_copy_f64(b.ptr, a.ptr, n, None))
    u.synchronize()
    tm = u.Timer()
    tm.start()
    for _ in range(reps):
        u.check(L.uhs_copy_f64(b.ptr, a.ptr, n, None))
    tm.stop()
    ms = tm.ms()
    a.close()
    b.close()
    return 2.0 * nbytes * reps / (ms * 1e-3) / 1e9


def batched(u, T=50, probe_T=8, tune=True, kinds=None):
    """64 uneven 3D grids (extents from {64,96,...,384}), half 3d1r_star and
    half 3d2r_box (125 points), seeded normalised coefficients, T steps: the
    plain per-grid sweeps against the batched graph.  The batch configuration
    is the learned cost model's pick: the model is trained on timings of
    random other batches, predicts every candidate for this batch, and the
    top 5 are timed on a probe_T-step run; the exhaustive timed search over
    all candidates is run beside it as the judge of that pick."""
    from uxxxxxxxxh_amd import costmodel as cm
    from uxxxxxxxxh_amd.workloads import uneven_batch
    specs = [g for g in uneven_batch() if not kinds or g["case"] in kinds]
    grids = cm.batch_grids(specs)
    pts = sum(int(np.prod(g["ext"])) for g in specs)

    def timed(fn, reps=2):
        fn()
        u.synchronize()
        tm = u.Timer()
        tm.start()
        for _ in range(reps):
            fn()
        tm.stop()
        return tm.ms() / reps

    model_info = None
    if tune:
        X, y, t_collect = cm.batch_collect()
        model, loss, t_train = cm.batch_train(X, y)
        ev = cm.batch_evaluate(model, grids, probe_T)
        model.close()
        best, ex_best = ev.pop("_model_best"), ev.pop("_exhaustive_best")
        model_info = dict(ev, training_rows=int(len(X)), collect_s=round(t_collect, 1),
                          train_s=round(t_train, 1), final_train_loss=round(loss, 4))
        log("batch cost model:", json.dumps(model_info))
    else:
        best = ex_best = u.batch_default_config()
    b0 = u.Batch(grids, T, u.batch_default_config())
    ms_plain = timed(b0.run_each)
    b0.close()

    def each_default():
        # Each grid through uhs_sweep_f64 with the library's per-grid default
        # (two steps per launch for the 7-point star grids).
        for A, B, ext, h, st in grids:
            u.sweep_f64(A, B, ext, h, st, T)
    ms_default = timed(each_default)
    b = u.Batch(grids, T, best)
    ms_batch = timed(b.run)
    info = b.info()
    b.close()
    if bytes(ex_best) != bytes(best):
        b = u.Batch(grids, T, ex_best)
        ms_ex = timed(b.run)
        b.close()
    else:
        ms_ex = ms_batch
    # Splitting the batch over 8 GPUs (a model; this run uses one GPU):
    # longest-processing-time-first packing on each grid's measured plain
    # sweep time, and on the library's estimate (points x work per point)
    # judged by the measured times.
    cost = []
    for A, B, ext, h, st in grids:
        cost.append(timed(lambda: u.sweep_f64(A, B, ext, h, st, 4), reps=1))
    gpu_m, load_m = u.batch_partition(grids, 8, cost)
    gpu_e, _ = u.batch_partition(grids, 8)
    load_e = [sum(c for c, g in zip(cost, gpu_e) if g == k) for k in range(8)]
    mean = sum(cost) / 8
    partition = {"gpus": 8, "cost": "each grid's plain sweep, 4 steps (ms)",
                 "max_over_mean_measured_costs": round(max(load_m) / mean, 4),
                 "max_over_mean_estimated_costs": round(max(load_e) / mean, 4),
                 "grids_per_gpu": [gpu_m.count(k) for k in range(8)]}
    cm.close_grids(grids)
    upd = float(pts) * T
    cfgd = lambda c: {"star_depth": c.star_depth, "star_tile": [c.star_tile_x, c.star_tile_y, c.star_tile_z],
                      "box_tile": [c.box_tile_x, c.box_tile_y, c.box_tile_z], "flags": c.flags}
    return {"grids": len(grids), "timesteps": T, "updates": upd,
            "plain_ms": round(ms_plain, 3), "plain_GStencils_s": round(upd / ms_plain / 1e6, 2),
            "batched_ms": round(ms_batch, 3), "batched_GStencils_s": round(upd / ms_batch / 1e6, 2),
            "speedup": round(ms_plain / ms_batch, 3),
            "tuned_by": "cost model top-5" if tune else "default configuration",
            "per_grid_default_ms": round(ms_default, 3),
            "speedup_vs_per_grid_default": round(ms_default / ms_batch, 3),
            "exhaustive_pick_batched_ms": round(ms_ex, 3),
            "speedup_exhaustive_pick": round(ms_plain / ms_ex, 3),
            "config": cfgd(best), "exhaustive_config": cfgd(ex_best),
            "cost_model": model_info,
            "launches": info["launches"], "tasks": info["tasks"], "partition": partition}


def batch_graph(u, T, kinds=None, runs=2):
    """The benchmark batch's graph alone (default configuration): two runs,
    each timed, with a progress line per run."""
    from uxxxxxxxxh_amd.workloads import uneven_batch
    grids, pts = [], 0
    for g in uneven_batch():
        if kinds and g["case"] not in kinds:
            continue
        st = u.builtin(g["case"])
        st = st.with_coeffs(u.make_coeffs(st, g["coeff_seed"]))
        h = st.radius
        A = u.DeviceArray(u.grid_shape(g["ext"], h), np.float64)
        B = u.DeviceArray(u.grid_shape(g["ext"], h), np.float64)
        u.fill_uniform(A, g["ext"], h, g["seed_a"])
        u.fill_uniform(B, g["ext"], h, g["seed_b"])
        grids.append((A, B, g["ext"], h, st))
        pts += int(np.prod(g["ext"]))
    b = u.Batch(grids, T, u.batch_default_config())
    log("batch graph", b.info())
    ms = []
    for r in range(runs):
        tm = u.Timer()
        tm.start()
        b.run()
        tm.stop()
        ms.append(tm.ms())
        log("batch graph run %d: %.3f ms" % (r, ms[-1]))
    b.close()
    for A, B, _, _, _ in grids:
        A.close()
        B.close()
    return {"grids": len(grids), "timesteps": T, "ms": ms, "GStencils_s": pts * T / (min(ms) * 1e-3) / 1e9}


# The reference's multi-kernel and one-sided example programs
# (UHStencil/examples/2d5pt_nested/stencil_2d5pt_nested.dsl,
# 3d7pt9pt_nested/stencil_3d7pt9pt_nested.dsl,
# 2d5pt_arbitrary_shape/stencil_2d5pt_arbitrary_shape.dsl and
# 3d7pt_arbitrary_shape/stencil_3d7pt_arbitrary_shape.dsl), at their own grid
# sizes, domains and iteration counts, as the DSL program reader reads them.
NESTED_PROGRAMS = {
    "2d5pt_nested": dict(
        dims=2, shape=[4098, 4098], lo=[17, 17], hi=[4081, 4081], T=20,
        producers=[([[-1, 0], [0, -1], [0, 0], [0, 1], [1, 0]], [0.2, 0.4, -0.5, 0.4, 0.2])],
        offsets=[[-1, -1], [-1, 1], [0, 0], [1, -1], [1, 1]], coeffs=[0.1, 0.3, -0.5, 0.3, 0.1],
        sources=[0, 0, 0, 0, 0]),
    "3d7pt9pt_nested": dict(
        dims=3, shape=[258, 258, 258], lo=[5, 5, 5], hi=[253, 253, 253], T=20, producers=[],
        offsets=[[-1, -1, -1], [-1, -1, 1], [-1, 1, -1], [-1, 1, 1], [0, 0, 0],
                 [1, -1, -1], [1, -1, 1], [1, 1, -1], [1, 1, 1]],
        coeffs=[0.1, 0.3, 0.7, 0.9, 0.5, 0.1, 0.3, 0.7, 0.9], sources=[-1] * 9),
    "2d5pt_arbitrary_shape": dict(
        dims=2, shape=[4098, 4098], lo=[2, 0], hi=[4098, 4096], T=20, producers=[],
        offsets=[[-2, 0], [-1, 0], [0, 0], [0, 1], [0, 2]], coeffs=[0.1, 0.2, 0.3, 0.4, 0.5], sources=[-1] * 5),
    "3d7pt_arbitrary_shape": dict(
        dims=3, shape=[258, 258, 258], lo=[0, 0, 0], hi=[256, 256, 256], T=20, producers=[],
        offsets=[[2, 0, 0], [1, 0, 0], [0, 0, 0], [0, 0, 1], [0, 0, 2], [0, 1, 0], [0, 2, 0]],
        coeffs=[0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7], sources=[-1] * 7),
}


def programs(u, reps=3, names=None):
    """Fused multi-kernel programs: the tiled kernel (the library's default
    tile, and the tile the program tuner picks) and the pointwise kernel, GStencils/s over the written domain, against
    16 algorithmic bytes per written cell and the FP64 rate of the fused
    formula (each producer re-evaluated at every access)."""
    out = {}
    for name, d in NESTED_PROGRAMS.items():
        if names and name not in names:
            continue
        prog = u.ProgramSpec(d["dims"], d["producers"], d["offsets"], d["coeffs"], d["sources"], d["lo"], d["hi"])
        A = u.DeviceArray(d["shape"], np.float64)
        B = u.DeviceArray(d["shape"], np.float64)
        ext = [n - 2 for n in d["shape"]]
        u.fill_uniform(A, ext, 1, 5)
        u.fill_uniform(B, ext, 1, 6)
        cells = float(np.prod([h - l for l, h in zip(d["lo"], d["hi"])]))
        flops = sum(2 * len(d["producers"][k][1]) - 1 if k >= 0 else 0 for k in d["sources"]) + 2 * len(d["sources"]) - 1
        roof = min(PEAK_HBM_GBPS / 16.0, PEAK_FP64_TFLOPS * 1e3 / flops)
        r = {"shape": d["shape"], "domain": [d["lo"], d["hi"]], "timesteps": d["T"],
             "flops_per_cell": flops, "roof_GStencils_s": round(roof, 1)}
        def tuned():
            # Tuning makes the winner the untiled call's choice, so it runs
            # after the default tile is timed.
            t0 = time.time()
            tile, _, n = u.tune_program_f64(A, B, d["shape"], prog, reps=3)
            r["tuning"] = {"tile": tile, "tiles_timed": n, "tuning_s": round(time.time() - t0, 3)}
            return tile
        for label, kernel, tile in (("tiled", u.PROGRAM_AUTO, None), ("tuned", u.PROGRAM_TILED, tuned),
                                    ("pointwise", u.PROGRAM_POINTWISE, None)):
            if callable(tile):
                tile = tile()
            u.program_sweep_f64(A, B, d["shape"], prog, 2, kernel=kernel, tile=tile)
            u.synchronize()
            tm = u.Timer()
            tm.start()
            for _ in range(reps):
                u.program_sweep_f64(A, B, d["shape"], prog, d["T"], kernel=kernel, tile=tile)
            tm.stop()
            ms = tm.ms() / (reps * d["T"])
            g = cells / (ms * 1e-3) / 1e9
            r[label] = {"ms_per_sweep": round(ms, 4), "GStencils_s": round(g, 2),
                        "hbm_frac_at_16B": round(16 * g / PEAK_HBM_GBPS, 4), "roof_frac": round(g / roof, 4)}
        r["tuned"].update(r.pop("tuning"))
        tuned_tile = r["tuned"]["tile"]
        # FP32 storage, Kahan-compensated sums (the DSL's float mode), default tile: 8 B per cell.
        A32 = u.DeviceArray(d["shape"], np.float32)
        B32 = u.DeviceArray(d["shape"], np.float32)
        u.fill_uniform(A32, ext, 1, 5)
        u.fill_uniform(B32, ext, 1, 6)
        u.program_sweep_f32(A32, B32, d["shape"], prog, 2, compensated=True)
        u.synchronize()
        tm = u.Timer()
        tm.start()
        for _ in range(reps):
            u.program_sweep_f32(A32, B32, d["shape"], prog, d["T"], compensated=True)
        tm.stop()
        ms = tm.ms() / (reps * d["T"])
        g = cells / (ms * 1e-3) / 1e9
        r["f32_kahan"] = {"ms_per_sweep": round(ms, 4), "GStencils_s": round(g, 2),
                          "hbm_frac_at_8B": round(8 * g / PEAK_HBM_GBPS, 4)}
        A32.close()
        B32.close()
        A.close()
        B.close()
        out[name] = r
        log("program %-16s tiled %.1f GSt/s, tuned %.1f GSt/s %s, pointwise %.1f GSt/s, f32 kahan %.1f GSt/s "
            "(roof %.0f)" % (name, r["tiled"]["GStencils_s"], r["tuned"]["GStencils_s"], tuned_tile,
                              r["pointwise"]["GStencils_s"], r["f32_kahan"]["GStencils_s"], roof))
    return out


def costmodel_leg(u, problems_per_shape=9, configs_per_problem=150, epochs=300, topk=5, seed=11):
    """The learned cost model of the per-grid tiling: timings of random
    problems of all 12 stencils (FP64, and FP32 plain and compensated; about
    20k timed configurations at the defaults),
    trained on the GPU, then judged on the two held-out problems (3d3r_box FP64
    384x512x640, 2d3r_star FP64 12288x6144) against the exhaustive timed
    search, the hand-set default and random-k."""
    import tempfile
    from uxxxxxxxxh_amd import costmodel as cm
    with tempfile.TemporaryDirectory() as d:
        db = os.path.join(d, "tune_db.csv")
        t0 = time.time()
        n = cm.collect(db, problems_per_shape, configs_per_problem, seed)
        t_collect = time.time() - t0
        X, y = cm.load_rows(db)
    idx = np.random.default_rng(seed).permutation(len(X))
    n_val = max(1, len(X) // 10)
    val, tr = idx[:n_val], idx[n_val:]
    # Three models from three seeds; configurations are ranked by the mean of
    # their predicted log times.
    models = [u.MLP(seed + i) for i in range(3)]
    t1 = time.time()
    for i, m in enumerate(models):
        m.train(X[tr], y[tr], epochs=epochs, batch=256, lr=1e-3, fit_norm=True, seed=seed + i)
    t_train = time.time() - t1
    pv = np.mean([m.predict(X[val]) for m in models], axis=0)
    rk = lambda v: np.argsort(np.argsort(v))
    out = {"timed_configs": n, "rows": int(len(X)), "collect_s": round(t_collect, 1), "train_s": round(t_train, 1),
           "ensemble": len(models),
           "val_rmse_log_ms": round(float(np.sqrt(np.mean((pv - y[val]) ** 2))), 4),
           "val_rank_corr": round(float(np.corrcoef(rk(pv), rk(y[val]))[0, 1]), 4),
           "held_out": cm.evaluate(models, topk, seed)}
    for m in models:
        m.close()
    return out


def fp32_kahan(u):
    from uxxxxxxxxh_amd.fp32study import study
    r = study("3d2r_star", [512, 512, 512], [10, 100, 1000])
    return {"workload": "3d2r_star FP32 (Kahan-compensated) 512^3 vs the FP64 sweep on identical inputs",
            "GStencils_s": round(r["f32k_GStencils_s"], 2), "hbm_frac_at_8B_per_update": round(r["f32k_hbm_frac"], 4),
            "errors": r["checkpoints"]}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ext, case="3d1r_star", budget_s=10.0, oracle_budget_s=5.0):
    """The oracle's restatement of the reference's serial sweep (cpu_verify,
    plain multiply-then-add in term order) on the same workload shape: one
    pinned thread (the baseline), then every host core this process may use
    (the OpenMP oracle, bitwise the same per point; reported beside it)."""
    from tests.oracle_py import Oracle, PLAIN
    import json as _json
    orc = Oracle()
    shapes = _json.load(open(os.path.join(ROOT, "tests", "golden", "shapes.json")))
    s = shapes[case]
    h = s["radius"]
    c = orc.make_coeffs(s["offsets"], s["dims"], 1)
    A = orc.fill(ext, h, 1)
    B = np.zeros_like(A)
    try:
        cpu = sorted(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        cpu = None

    def run(nthreads, budget):
        sweeps, t0 = 0, time.perf_counter()
        while True:
            src, dst = (A, B) if sweeps % 2 == 0 else (B, A)
            orc.sweep(src, dst, ext, h, s["offsets"], c, PLAIN, nthreads=nthreads)
            sweeps += 1
            el = time.perf_counter() - t0
            if el >= budget:
                return sweeps, el

    if cpu:
        os.sched_setaffinity(0, {cpu[0]})
    sweeps, el = run(1, budget_s)
    if cpu:
        os.sched_setaffinity(0, set(cpu))
    n = float(np.prod(ext))
    threads = len(cpu) if cpu else (os.cpu_count() or 1)
    osw, oel = run(threads, oracle_budget_s)
    return {"value": n * sweeps / el / 1e9, "unit": "GStencils/s", "cores": 1, "kind": "port",
            "sample": "%s FP64 %s, %d sweeps (%.1f s), oracle plain mode, 1 pinned thread"
                      % (case, "x".join(map(str, ext)), sweeps, el),
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(), "cpus_usable": threads,
            "oracle_all_cores": {"value": n * osw / oel / 1e9, "unit": "GStencils/s", "threads": threads,
                                 "sample": "%d sweeps (%.1f s), same grid, OpenMP over rows" % (osw, oel),
                                 "label": "oracle (not the baseline)"}}


def kernel_name(cfg):
    if cfg.flags & 32:
        return "k_star3d_tb2<double,Prog_3d1r_star,%d,false,%d>" % (cfg.vec, 2 if cfg.flags & 2 else 1)
    if cfg.flags & 4:
        return "k_star3d_rt<double,Prog_3d1r_star,%d,false>" % cfg.vec
    nt = cfg.tile_x // cfg.vec * cfg.tile_y
    ring = 2 * (cfg.tile_x + cfg.tile_y)
    nh = -(-ring // nt)
    nh = 1 if nh <= 1 else 2 if nh <= 2 else 4
    return "k_star3d<double,Prog_3d1r_star,%d,false,%d,%d>" % (cfg.vec, 2 if cfg.flags & 2 else 1, nh)


def load_traffic(case, ext):
    p = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (case, "x".join(map(str, ext))))
    if os.path.exists(p):
        with open(p) as fh:
            return json.load(fh)
    return None


def probes(u, out):
    """The copy-bandwidth and FP64 FMA probes: this device's reachable rates."""
    try:
        out["measured_copy_GBps"] = round(copy_bandwidth(u), 1)
        log("device copy bandwidth %.0f GB/s" % out["measured_copy_GBps"])
    except Exception as e:  # context only
        log("copy bandwidth probe failed:", e)
    try:
        out["measured_fp64_TFLOPs"] = round(fma_rate(u), 2)
        log("device FP64 FMA rate %.1f TFLOP/s" % out["measured_fp64_TFLOPs"])
    except Exception as e:  # context only
        log("FMA rate probe failed:", e)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--timesteps", type=int, default=100)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--no-shapes", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-costmodel", action="store_true")
    ap.add_argument("--costmodel-only", action="store_true", help="only the per-grid cost-model leg")
    ap.add_argument("--no-one-step", action="store_true")
    ap.add_argument("--one-step-only", action="store_true",
                    help="only the one-step-per-launch sweep (its traffic-counter run)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline sweep (profiling runs: no probes, legs or CPU baseline)")
    ap.add_argument("--no-programs", action="store_true")
    ap.add_argument("--shape-only", default=None, metavar="CASE",
                    help="only one benchmark stencil at its per-shape size (copy probe included)")
    ap.add_argument("--no-shard", action="store_true",
                    help="skip one GPU's share of the slab decomposition")
    ap.add_argument("--programs-only", nargs="*", default=None, metavar="NAME",
                    help="only the fused multi-kernel programs (all, or those named)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched leg (profiling runs)")
    ap.add_argument("--batch-graph-only", action="store_true",
                    help="only the batched graph, default configuration, --timesteps steps, two runs "
                         "(counter runs: few dispatches, no per-grid legs)")
    ap.add_argument("--no-tune", action="store_true",
                    help="batched leg with the default batch configuration (no tuning search)")
    ap.add_argument("--batch-kinds", default=None,
                    help="batched leg on the grids of these stencils only (comma-separated)")
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32", "f32k"],
                    help="precision of --shape-only")
    ap.add_argument("--copy-only", action="store_true",
                    help="only the copy kernel (traffic-counter calibration: 2 GiB read + written)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import uxxxxxxxxh_amd as u
    u.lib()
    if args.copy_only:
        print(json.dumps({"copy_GBps": round(copy_bandwidth(u), 1)}))
        return
    if args.shape_only:
        # One benchmark stencil at its per-shape size (profiling runs).
        case = args.shape_only
        e = [8192, 8192] if case.startswith("2d") else [512, 512, 512]
        ms_s, ms_w, c = time_sweeps(u, case, e, args.timesteps, args.steps, args.warmup, dtype=args.dtype)
        print(json.dumps({"case": case, "GStencils_s": round(float(np.prod(e)) / (ms_w * 1e-3) / 1e9, 2),
                          "ms_per_sweep": round(ms_w, 5), "tile": list(c.tuple()),
                          "copy_GBps": round(copy_bandwidth(u, reps=2), 1)}))
        return
    if args.costmodel_only:
        print(json.dumps(costmodel_leg(u)))
        return
    if args.batch_graph_only:
        kinds = args.batch_kinds.split(",") if args.batch_kinds else None
        print(json.dumps(batch_graph(u, args.timesteps, kinds)))
        return
    if args.batch_only:
        kinds = args.batch_kinds.split(",") if args.batch_kinds else None
        print(json.dumps(batched(u, tune=not args.no_tune, kinds=kinds)))
        return
    if args.programs_only is not None:
        print(json.dumps(programs(u, names=args.programs_only)))
        return
    if args.one_step_only:
        print(json.dumps(one_step(u, args.size, args.timesteps)))
        return
    if args.headline_only:
        for k in ("no_shapes", "no_one_step", "no_programs", "no_shard", "no_batch", "no_fp32", "no_cpu",
                  "no_probes", "no_costmodel"):
            setattr(args, k, True)
    name = __import__("ctypes").create_string_buffer(256)
    u.check(u.lib().uhs_device_name(name, 256))
    log("device:", name.value.decode())

    n = args.size
    ext = [n, n, n]
    T = args.timesteps
    ms_step, ms_sweep, cfg = time_sweeps(u, "3d1r_star", ext, T, args.steps, args.warmup)
    updates = float(n) ** 3
    gst = updates * T / (ms_step * 1e-3) / 1e9
    # The default advances two steps per launch (UHS_CFG_TEMPORAL2) at this
    # size; T = 100 is even, so every launch is a two-step one.
    per_launch = 2 if cfg.flags & 32 else 1
    ms_launch = ms_sweep * per_launch
    alg_bytes = 16.0 * updates * per_launch
    achieved = alg_bytes / (ms_launch * 1e-3) / 1e9
    log("3d1r_star %d^3: %.3f ms/launch, %.1f GSt/s, %.0f GB/s, cfg %s" % (n, ms_launch, gst, achieved, cfg))
    traffic = load_traffic("3d1r_star_tb2" if per_launch == 2 else "3d1r_star", ext)
    out = {
        "metric": "GStencils/s (FP64 3d1r_star 512^3 sweep, T=100)",
        "value": round(gst, 3),
        "unit": "GStencils/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # Context, not a baseline on this hardware: the reference's own
        # published per-sweep time for this stencil and size on NVIDIA A100
        # (UHStencil's figure data, 1.676 ms per 512^3 sweep).
        "reference_other_hw": {"hardware": "NVIDIA A100", "GStencils_s": 80.1},
        "dtype": "f64",
        "data": "synthetic (seeded U[0,1) interior, zero halo, symmetric normalised coefficients)",
        "config": {"workload": "3d1r_star FP64 %dx%dx%d interior, halo 1, T=%d per step" % (n, n, n, T),
                   "tile": [cfg.tile_x, cfg.tile_y, cfg.tile_z, cfg.vec, cfg.flags]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBPS, 4),
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "kernel": kernel_name(cfg),
                     "steps_per_launch": per_launch,
                     "algorithmic_bytes_per_launch": int(alg_bytes),
                     "avg_launch_ms": round(ms_launch, 5)},
    }
    if traffic:
        # The DRAM bytes the launch really moves (traffic counters, committed
        # summary) over the same launch time: with two steps per launch the
        # kernel moves fewer bytes than the 16 per update counted above.
        dram = traffic["bytes_per_launch"] / (ms_launch * 1e-3) / 1e9
        out["roofline"]["dram_GBps"] = round(dram, 1)
        out["roofline"]["dram_frac"] = round(dram / PEAK_HBM_GBPS, 4)
        out["roofline"]["dram_bytes_per_update"] = round(traffic["bytes_per_launch"] / (updates * per_launch), 3)
    if not getattr(args, "no_probes", False):
        probes(u, out)
    if not args.no_shapes:
        per = {}
        for case in u.builtin_names():
            dims = 2 if case.startswith("2d") else 3
            e = [8192, 8192] if dims == 2 else [512, 512, 512]
            ms_s, ms_w, c = time_sweeps(u, case, e, T, 2, 1)
            pts = float(np.prod(e))
            st = u.builtin(case)
            flops = 2 * st.npoints - 1
            roof = min(PEAK_HBM_GBPS / 16.0, PEAK_FP64_TFLOPS * 1e3 / flops)
            g = pts / (ms_w * 1e-3) / 1e9
            per[case] = {"GStencils_s": round(g, 2), "ms_per_sweep": round(ms_w, 4),
                         "hbm_frac": round(g * 16 / PEAK_HBM_GBPS, 4),
                         "roof_GStencils_s": round(roof, 1), "roof_frac": round(g / roof, 4),
                         "tile": [c.tile_x, c.tile_y, c.tile_z, c.vec, c.flags]}
            if "measured_copy_GBps" in out and "measured_fp64_TFLOPs" in out:
                # The same roof from the rates this device reaches (copy kernel,
                # FMA probe) instead of the published peaks.
                mroof = min(out["measured_copy_GBps"] / 16.0, out["measured_fp64_TFLOPs"] * 1e3 / flops)
                per[case]["measured_roof_GStencils_s"] = round(mroof, 1)
                per[case]["measured_roof_frac"] = round(g / mroof, 4)
            log("%-10s %8.2f GSt/s  %.3f ms/sweep  roof %.0f  frac %.3f" % (case, g, ms_w, roof, g / roof))
        out["per_shape"] = per
    if not args.no_one_step:
        out["one_step"] = one_step(u, n, T)
        log("one step per launch", out["one_step"])
    if not args.no_programs:
        out["programs"] = programs(u)
    if not args.no_shard:
        out["shard_share"] = shard_share(u, T)
        log("shard share", out["shard_share"])
    if not args.no_batch:
        out["batched"] = batched(u)
        log("batched", out["batched"])
    if not args.no_costmodel:
        out["costmodel"] = costmodel_leg(u)
        log("costmodel", json.dumps(out["costmodel"]))
    if not args.no_fp32:
        out["fp32_kahan"] = fp32_kahan(u)
        log("fp32", out["fp32_kahan"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ext)
        log("cpu baseline", out["cpu_baseline"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
