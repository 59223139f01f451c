"""FP32 storage with Kahan-compensated neighbour sums, against the FP64 sweep.

    python -m uxxxxxxxxh_amd.fp32study [--cases 3d2r_star,...] [--steps 10,100,1000]

For each stencil: identical seeded inputs (U[0,1) interior, zero halo, the
FP32 grid is the FP64 grid rounded) and normalised symmetric coefficients;
the FP64 form (bitwise equal to the CPU oracle's fused sweep), the plain FP32
form and the compensated FP32 form are advanced side by side, and at each
checkpoint the error of each FP32 form against FP64 is reduced on the GPU in
FP64: max |a - b| and relative L2 = ||a - b|| / ||b||.  Also times the
compensated form (GStencils/s) on the same grid.
"""
import argparse
import json
import sys

import numpy as np

from . import (DeviceArray, Timer, builtin, builtin_names, error_norms, fill_uniform, grid_shape, lib,
               make_coeffs, sweep_f32, sweep_f64, synchronize)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def study(case, ext, checkpoints, seed=3):
    st = builtin(case)
    st = st.with_coeffs(make_coeffs(st, seed))
    h = st.radius
    shape = grid_shape(ext, h)
    g = {}
    for name, dt in (("f64", np.float64), ("f32", np.float32), ("f32k", np.float32)):
        A, B = DeviceArray(shape, dt), DeviceArray(shape, dt)
        fill_uniform(A, ext, h, seed)
        fill_uniform(B, ext, h, seed)
        g[name] = (A, B)
    out = {"extents": list(ext), "checkpoints": {}}
    done = 0
    for cp in checkpoints:
        n = cp - done
        assert n % 2 == 0, "checkpoints keep the state in A (even step counts)"
        sweep_f64(*g["f64"], ext, h, st, n)
        sweep_f32(*g["f32"], ext, h, st, n, compensated=False)
        sweep_f32(*g["f32k"], ext, h, st, n, compensated=True)
        done = cp
        row = {}
        for name in ("f32", "f32k"):
            mx, d2, r2 = error_norms(g[name][0], g["f64"][0], ext, h)
            row[name] = {"max_abs": mx, "rel_l2": d2 / r2 if r2 else 0.0}
        out["checkpoints"][str(cp)] = row
        log("%-10s step %5d  plain fp32: max %.3e rel-L2 %.3e | kahan fp32: max %.3e rel-L2 %.3e"
            % (case, cp, row["f32"]["max_abs"], row["f32"]["rel_l2"], row["f32k"]["max_abs"], row["f32k"]["rel_l2"]))
    # throughput of the compensated form
    tm = Timer()
    reps = 20
    sweep_f32(*g["f32k"], ext, h, st, 2, compensated=True)
    synchronize()
    tm.start()
    sweep_f32(*g["f32k"], ext, h, st, reps, compensated=True)
    tm.stop()
    ms = tm.ms() / reps
    out["f32k_ms_per_sweep"] = ms
    out["f32k_GStencils_s"] = float(np.prod(ext)) / (ms * 1e-3) / 1e9
    out["f32k_hbm_frac"] = out["f32k_GStencils_s"] * 8 / 8000.0
    for A, B in g.values():
        A.close()
        B.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", default="3d2r_star")
    ap.add_argument("--steps", default="10,100,1000")
    ap.add_argument("--size3d", type=int, default=512)
    ap.add_argument("--size2d", type=int, default=8192)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    lib()
    cases = builtin_names() if a.cases == "all" else a.cases.split(",")
    cps = [int(v) for v in a.steps.split(",")]
    res = {}
    for case in cases:
        ext = [a.size2d] * 2 if case.startswith("2d") else [a.size3d] * 3
        res[case] = study(case, ext, cps)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(res, fh, indent=1)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
