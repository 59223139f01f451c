"""Parity of the HIP sweep kernels (through the C ABI) with the CPU oracle.

Bar: bitwise equality with the oracle's fused mode (the kernels' exact
operation sequence), and <= 1e-12 relative against its plain mode (the
reference's separate multiply and add), on the same seeded inputs, for all
12 benchmark stencils, partial edge tiles, random (nonzero) halos and
T = 1, 2, 3 ping-pong."""
import json
import os
import sys

import numpy as np
import pytest

from tests.oracle_py import FUSED, KAHAN, PLAIN

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = json.load(open(os.path.join(HERE, "golden", "shapes.json")))
CASES = list(SHAPES)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def small_ext(dims):
    return [67, 301] if dims == 2 else [19, 37, 70]


def rel_err(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def run_gpu(u, st, A0, B0, ext, h, T, dtype=np.float64, compensated=False, cfg=None):
    A = u.DeviceArray.from_host(A0.astype(dtype))
    B = u.DeviceArray.from_host(B0.astype(dtype))
    try:
        if dtype == np.float64:
            u.sweep_f64(A, B, ext, h, st, T, cfg)
        else:
            u.sweep_f32(A, B, ext, h, st, T, compensated, cfg)
        u.synchronize()
        return A.download(), B.download()
    finally:
        A.close()
        B.close()


def random_grids(ext, h, seed):
    rng = np.random.default_rng(seed)
    shape = [e + 2 * h for e in ext]
    return rng.random(shape), rng.random(shape)


def test_fill_matches_oracle(gpu, orc):
    u = gpu
    for ext, h in (([33, 40, 75], 2), ([100, 257], 3)):
        for dt in (np.float64, np.float32):
            d = u.DeviceArray([e + 2 * h for e in ext], dt)
            u.fill_uniform(d, ext, h, 99)
            got = d.download()
            d.close()
            assert np.array_equal(got, orc.fill(ext, h, 99, dt))


def last_launch_steps(u, flags, T, K=2):
    """Steps the last launch of a T-step sweep advances: the other grid then
    holds state T minus that.  With K steps per launch (UHS_CFG_TEMPORAL2:
    K = 2; UHS_CFG_TEMPORALK: K = tile_y) the plan runs T / K K-step
    launches, for even K an even number of them, after the leading steps:
    launches of 5 and then 3 steps where they are shorter than K (the
    defaults here have one column per thread, so both compile), then single
    sweeps."""
    if not flags & (u.CFG_TEMPORAL2 | u.CFG_TEMPORALK):
        return 1
    if not flags & u.CFG_TEMPORALK:
        K = 2
    launches, ones = T // K, T % K
    if K % 2 == 0 and launches % 2:
        launches -= 1
        ones += K
    seq = []
    if flags & u.CFG_TEMPORALK:
        for k in (5, 3):
            while k < K and ones >= k:
                seq.append(k)
                ones -= k
    seq += [1] * ones + [K] * launches
    return seq[-1] if seq else 1


@pytest.mark.parametrize("case", CASES)
def test_f64_parity_all_T(gpu, orc, case):
    """The default configuration: the result grid bitwise the oracle's FUSED
    mode and within 1e-12 of its PLAIN mode, for T = 1..4; the other grid
    holds the state before the last launch (T-1 after a single step, T-2 or
    T-3 after a multi-step launch)."""
    u = gpu
    s = SHAPES[case]
    st = u.builtin(case)
    c = u.make_coeffs(st, 2024)
    st = st.with_coeffs(c)
    ext, h = small_ext(s["dims"]), s["radius"]
    A0, B0 = random_grids(ext, h, 7)
    dcfg = u.default_config(ext, h, st)
    flags = dcfg.flags
    inner = tuple(slice(h, h + e) for e in ext)
    log(case, "kernel", u.plan_kind(ext, h, st), "ext", ext, "flags", flags)
    for T in (1, 2, 3, 4):
        gA, gB = run_gpu(u, st, A0, B0, ext, h, T)
        got, other = (gA, gB) if T % 2 == 0 else (gB, gA)
        Tp = T - last_launch_steps(u, flags, T, dcfg.tile_y)
        for mode in (FUSED, PLAIN):
            # states[k]: the oracle's state after k single sweeps
            states = {0: None}
            A, B = A0.copy(), B0.copy()
            for k in range(1, T + 1):
                if k % 2:
                    orc.iterate(A, B, ext, h, s["offsets"], c, 1, mode)
                else:
                    orc.iterate(B, A, ext, h, s["offsets"], c, 1, mode)
                states[k] = (B if k % 2 else A).copy()
            want = states[T]
            # The other grid keeps its own halo; its interior holds state Tp
            # (its initial content before any launch wrote it).
            other0 = B0 if T % 2 == 0 else A0
            want_other = other0.copy()
            if Tp > 0:
                want_other[inner] = states[Tp][inner]
            if mode == FUSED:
                assert np.array_equal(got, want), (case, T)
                assert np.array_equal(other, want_other), (case, T, Tp)
            else:
                assert rel_err(got, want) <= 1e-12 and rel_err(other, want_other) <= 1e-12


@pytest.mark.parametrize("case", CASES)
def test_dsl_coefficients_and_halo_untouched(gpu, orc, case):
    u = gpu
    s = SHAPES[case]
    st = u.builtin(case)  # the DSL's own coefficients
    ext, h = small_ext(s["dims"]), s["radius"]
    A0, B0 = random_grids(ext, h, 8)
    gA, gB = run_gpu(u, st, A0, B0, ext, h, 1)
    A, B = A0.copy(), B0.copy()
    orc.iterate(A, B, ext, h, s["offsets"], s["coeffs"], 1, FUSED)
    assert np.array_equal(gB, B) and np.array_equal(gA, A0)
    mask = np.ones(B0.shape, bool)
    mask[tuple(slice(h, n - h) for n in B0.shape)] = False
    assert np.array_equal(gB[mask], B0[mask])


CONFIGS = {
    "star3d": [(64, 8, 64, 2), (32, 4, 16, 1), (128, 4, 8, 4), (64, 2, 5, 2), (256, 2, 3, 4),
               (128, 16, 9, 4, 2), (64, 8, 6, 2, 3), (256, 8, 7, 8, 9), (512, 4, 33, 8, 1), (64, 8, 5, 4, 8),
               # register-tiled kernel: 64-lane bands, vec rows per lane
               (64, 8, 16, 4, 4), (64, 4, 7, 2, 5), (64, 32, 64, 8, 4), (64, 16, 3, 8, 4), (64, 2, 11, 2, 4)],
    "box3d": [(64, 4, 128, 1), (32, 8, 7, 2), (16, 4, 3, 1), (128, 2, 11, 2), (32, 16, 5, 1, 1),
              (32, 16, 9, 1, 2), (64, 8, 6, 2, 3),
              # two rows per thread
              (32, 16, 9, 1, 16), (64, 8, 5, 1, 17), (16, 4, 3, 1, 16),
              # two planes of tile loads in flight
              (32, 16, 9, 1, 64), (32, 8, 7, 2, 65), (32, 16, 9, 1, 80), (16, 4, 3, 1, 80)],
    "win2d": [(256, 1, 64, 1), (128, 1, 5, 2), (64, 1, 17, 4), (512, 1, 9, 1), (1024, 1, 3, 2, 1),
              # two / four rows of loads in flight
              (256, 1, 5, 1, 2), (128, 1, 3, 2, 3), (64, 1, 7, 4, 8), (512, 1, 64, 1, 9),
              # ring of 2r+1 rows (whole ring turns plus a remainder)
              (256, 1, 5, 1, 128), (128, 1, 3, 2, 129), (64, 1, 17, 4, 128), (512, 1, 64, 1, 129)],
}


@pytest.mark.parametrize("case", CASES)
def test_config_invariance(gpu, orc, case):
    u = gpu
    s = SHAPES[case]
    st = u.builtin(case).with_coeffs(u.make_coeffs(u.builtin(case), 5))
    ext, h = small_ext(s["dims"]), s["radius"]
    A0, B0 = random_grids(ext, h, 9)
    A, B = A0.copy(), B0.copy()
    orc.iterate(A, B, ext, h, s["offsets"], st.coeffs, 2, FUSED)
    kind = u.plan_kind(ext, h, st)
    ran = 0
    for cfg in CONFIGS[kind]:
        try:
            u.plan_kind(ext, h, st, cfg=cfg)
        except u.UhsError as e:
            assert e.code == -3
            continue
        gA, gB = run_gpu(u, st, A0, B0, ext, h, 2, cfg=cfg)
        assert np.array_equal(gA, A) and np.array_equal(gB, B), (case, cfg)
        ran += 1
    assert ran >= 2


@pytest.mark.parametrize("case", ["3d1r_star", "3d2r_star", "3d3r_star"])
@pytest.mark.parametrize("dtype", ["f64", "f32", "f32k"])
def test_register_tiled_star_all_precisions(gpu, orc, case, dtype):
    u = gpu
    s = SHAPES[case]
    st = u.builtin(case).with_coeffs(u.make_coeffs(u.builtin(case), 12))
    ext, h = [13, 29, 131], s["radius"] + 1
    A0, B0 = random_grids(ext, h, 13)
    dt = np.float64 if dtype == "f64" else np.float32
    A0, B0 = A0.astype(dt), B0.astype(dt)
    for cfg in ((64, 8, 5, 4, 4), (64, 16, 64, 2, 5)):
        gA, gB = run_gpu(u, st, A0, B0, ext, h, 3, dt, dtype == "f32k", cfg=cfg)
        A, B = A0.copy(), B0.copy()
        orc.iterate(A, B, ext, h, s["offsets"], st.coeffs, 3, KAHAN if dtype == "f32k" else FUSED)
        assert np.array_equal(gA, A) and np.array_equal(gB, B), cfg


@pytest.mark.parametrize("case", ["3d1r_box", "3d2r_box", "3d3r_box"])
@pytest.mark.parametrize("dtype", ["f64", "f32", "f32k"])
def test_box_two_rows_all_precisions(gpu, orc, case, dtype):
    u = gpu
    s = SHAPES[case]
    st = u.builtin(case).with_coeffs(u.make_coeffs(u.builtin(case), 14))
    ext, h = [11, 27, 45], s["radius"]
    A0, B0 = random_grids(ext, h, 15)
    dt = np.float64 if dtype == "f64" else np.float32
    A0, B0 = A0.astype(dt), B0.astype(dt)
    for cfg in ((32, 16, 4, 1, 16), (16, 16, 64, 1, 17), (64, 8, 5, 1, 80), (64, 8, 3, 2, 64)):
        gA, gB = run_gpu(u, st, A0, B0, ext, h, 3, dt, dtype == "f32k", cfg=cfg)
        A, B = A0.copy(), B0.copy()
        orc.iterate(A, B, ext, h, s["offsets"], st.coeffs, 3, KAHAN if dtype == "f32k" else FUSED)
        assert np.array_equal(gA, A) and np.array_equal(gB, B), cfg


@pytest.mark.parametrize("case", ["3d1r_star", "3d2r_star", "3d3r_star"])
@pytest.mark.parametrize("dtype", ["f64", "f32", "f32k"])
def test_two_steps_per_launch(gpu, orc, case, dtype):
    """UHS_CFG_TEMPORAL2: the result grid (A for even T, B for odd T) is
    bitwise the oracle's; the halos of both grids, which differ and are
    nonzero, are untouched; every state's halo is read from the grid the
    one-step ping-pong holds it in.  T = 1..6 covers 0-3 leading single
    steps and 0-2 two-step launches."""
    u = gpu
    s = SHAPES[case]
    R = s["radius"]
    st = u.builtin(case).with_coeffs(u.make_coeffs(u.builtin(case), 16))
    dt = np.float64 if dtype == "f64" else np.float32
    T2, NT = u.CFG_TEMPORAL2, u.CFG_NT_STORE
    for ext, h in (([19, 37, 70], R), ([9, 14, 23], R + 1)):
        A0, B0 = random_grids(ext, h, 17)
        A0, B0 = A0.astype(dt), B0.astype(dt)
        mask = np.ones(A0.shape, bool)
        mask[tuple(slice(h, n - h) for n in A0.shape)] = False
        for cfg in ((10, 4, 5, 2, T2), (30, 6, 7, 1, T2 | NT), (62, 6, 64, 2, T2), (14, 6, 3, 2, T2),
                    (30, 4, 64, 1, T2 | NT)):
            for T in (1, 2, 3, 4, 5, 6):
                log(case, dtype, ext, h, cfg, "T", T)
                gA, gB = run_gpu(u, st, A0, B0, ext, h, T, dt, dtype == "f32k", cfg=cfg)
                A, B = A0.copy(), B0.copy()
                orc.iterate(A, B, ext, h, s["offsets"], st.coeffs, T, KAHAN if dtype == "f32k" else FUSED)
                got, want = (gA, A) if T % 2 == 0 else (gB, B)
                assert np.array_equal(got, want), (ext, h, cfg, T)
                assert np.array_equal(gA[mask], A0[mask]) and np.array_equal(gB[mask], B0[mask])


@pytest.mark.parametrize("dtype", ["f64", "f32", "f32k"])
def test_two_steps_per_launch_box3d(gpu, orc, dtype):
    """UHS_CFG_TEMPORAL2 for the radius-1 3D box: running sums per pending
    plane for both levels.  The result grid bitwise the oracle's for T =
    1..6 and 9, halos untouched and every state's halo read from the grid
    that holds it, on grids with interior and edge tiles and halos wider
    than the radius."""
    u = gpu
    case = "3d1r_box"
    s = SHAPES[case]
    st = u.builtin(case).with_coeffs(u.make_coeffs(u.builtin(case), 16))
    dt = np.float64 if dtype == "f64" else np.float32
    T2, NT = u.CFG_TEMPORAL2, u.CFG_NT_STORE
    for ext, h in (([19, 37, 70], 1), ([9, 14, 23], 2), ([40, 44, 70], 1)):
        A0, B0 = random_grids(ext, h, 5)
        A0, B0 = A0.astype(dt), B0.astype(dt)
        mask = np.ones(A0.shape, bool)
        mask[tuple(slice(h, n - h) for n in A0.shape)] = False
        for cfg in ((10, 4, 5, 2, T2), (30, 6, 7, 1, T2 | NT), (62, 6, 64, 2, T2), (14, 6, 3, 2, T2),
                    (30, 14, 9, 2, T2 | NT), (14, 6, 8, 4, T2)):
            if cfg[3] == 4 and dtype == "f64":
                with pytest.raises(u.UhsError):
                    u.plan_kind(ext, h, st, u.DT_F64, cfg=cfg)
                continue
            for T in (1, 2, 3, 4, 5, 6, 9):
                log(case, dtype, ext, h, cfg, "T", T)
                gA, gB = run_gpu(u, st, A0, B0, ext, h, T, dt, dtype == "f32k", cfg=cfg)
                A, B = A0.copy(), B0.copy()
                orc.iterate(A, B, ext, h, s["offsets"], st.coeffs, T, KAHAN if dtype == "f32k" else FUSED)
                got, want = (gA, A) if T % 2 == 0 else (gB, B)
                assert np.array_equal(got, want), (ext, h, cfg, T)
                assert np.array_equal(gA[mask], A0[mask]) and np.array_equal(gB[mask], B0[mask])


def two_step_forms_2d(case, dtype):
    """(form flag, columns per thread) the 2D two-step kernel is compiled
    for: without UHS_CFG_ROWRING (flag 0) the ring form of the stars at 1, 2
    and 4 columns and the sum form of the boxes at 1 and 2; the window form
    (UHS_CFG_ROWRING) where its two windows fit the registers."""
    R = SHAPES[case]["radius"]
    star = "star" in case
    f64 = dtype == "f64"
    win = [1, 2, 4] if R == 1 else ([1] if f64 else [1, 2]) if R == 2 else ([] if f64 and not star else [1])
    out = [(128, v) for v in win]
    out += [(0, v) for v in ((1, 2, 4) if star else (1, 2))]
    return out


@pytest.mark.parametrize("case", [c for c in CASES if c.startswith("2d")])
@pytest.mark.parametrize("dtype", ["f64", "f32", "f32k"])
def test_two_steps_per_launch_2d(gpu, orc, case, dtype):
    """UHS_CFG_TEMPORAL2 for the 2D stencils, both forms, as for the 3D
    stars: the result grid is bitwise the oracle's for T = 1..6 and 9, halos
    untouched and every state's halo read from the grid that holds it, with
    edge tiles in both axes, slabs of 3 to 64 rows and halos wider than the
    radius."""
    u = gpu
    s = SHAPES[case]
    R = s["radius"]
    st = u.builtin(case).with_coeffs(u.make_coeffs(u.builtin(case), 16))
    dt = np.float64 if dtype == "f64" else np.float32
    T2, NT = u.CFG_TEMPORAL2, u.CFG_NT_STORE
    forms = two_step_forms_2d(case, dtype)
    if not forms:
        with pytest.raises(u.UhsError):
            u.plan_kind([67, 301], R, st, u.DT_F64, cfg=(64 - 2 * R, 1, 16, 1, T2 | u.CFG_ROWRING))
        return
    for ext, h in (([67, 301], R), ([23, 45], R + 1), ([150, 40], R)):
        A0, B0 = random_grids(ext, h, 5)
        A0, B0 = A0.astype(dt), B0.astype(dt)
        mask = np.ones(A0.shape, bool)
        mask[tuple(slice(h, n - h) for n in A0.shape)] = False
        for form, v in forms:
            for bx, tz, fl in ((64, 16, T2), (16, 3, T2 | NT), (128, 5, T2), (256, 64, T2 | NT)):
                if bx % v or bx // v < 2 * R:  # the ring needs 2r staging threads
                    continue
                cfg = (bx - 2 * R, 1, tz, v, fl | form)
                for T in (1, 2, 3, 4, 5, 6, 9):
                    log(case, dtype, ext, h, cfg, "T", T)
                    gA, gB = run_gpu(u, st, A0, B0, ext, h, T, dt, dtype == "f32k", cfg=cfg)
                    A, B = A0.copy(), B0.copy()
                    orc.iterate(A, B, ext, h, s["offsets"], st.coeffs, T, KAHAN if dtype == "f32k" else FUSED)
                    got, want = (gA, A) if T % 2 == 0 else (gB, B)
                    assert np.array_equal(got, want), (ext, h, cfg, T)
                    assert np.array_equal(gA[mask], A0[mask]) and np.array_equal(gB[mask], B0[mask])


def test_two_steps_per_launch_config_checks_2d(gpu):
    u = gpu
    T2, RR = u.CFG_TEMPORAL2, u.CFG_ROWRING
    ext = [64, 96]
    s1, s3, b1 = u.builtin("2d1r_star"), u.builtin("2d3r_star"), u.builtin("2d1r_box")
    assert u.plan_kind(ext, 1, s1, cfg=(254, 1, 32, 4, T2)) == "win2d"
    assert u.plan_kind(ext, 1, s1, cfg=(254, 1, 32, 4, T2 | RR)) == "win2d"
    assert u.plan_kind(ext, 3, s3, cfg=(58, 1, 32, 1, T2 | RR)) == "win2d"
    assert u.plan_kind(ext, 3, s3, cfg=(58, 1, 32, 4, T2)) == "win2d"
    assert u.plan_kind(ext, 1, b1, cfg=(62, 1, 32, 2, T2 | RR)) == "win2d"
    assert u.plan_kind(ext, 1, b1, cfg=(62, 1, 32, 2, T2)) == "win2d"        # sum form
    assert u.plan_kind(ext, 3, u.builtin("2d3r_box"), cfg=(58, 1, 32, 1, T2)) == "win2d"
    for st, bad in ((s1, (61, 1, 32, 2, T2)),                  # tile_x + 2r odd with vec 2
                    (s1, (62, 2, 32, 2, T2)),                  # tile_y 2
                    (s1, (1022, 1, 32, 1, T2)),                # 1024 threads
                    (s1, (62, 1, 32, 2, T2 | u.CFG_PREFETCH2)),
                    (s3, (58, 1, 32, 2, T2 | RR)),             # FP64 radius-3 window: one column per thread
                    (b1, (62, 1, 32, 4, T2)),                  # sum form: 1 or 2 columns per thread
                    (u.builtin("2d3r_box"), (58, 1, 32, 1, T2 | RR))):  # no FP64 radius-3 box window form
        with pytest.raises(u.UhsError) as e:
            u.plan_kind(ext, st.radius, st, cfg=bad)
        assert e.value.code == -3, bad
    assert u.plan_kind(ext, 3, u.builtin("2d3r_box"), u.DT_F32, cfg=(58, 1, 32, 1, T2 | RR)) == "win2d"


@pytest.mark.parametrize("case", ["2d1r_star", "2d2r_star", "2d3r_star", "2d1r_box", "2d2r_box"])
@pytest.mark.parametrize("dtype", ["f64", "f32", "f32k"])
@pytest.mark.parametrize("K", [3, 4, 5, 6, 8])
def test_k_steps_per_launch_2d(gpu, orc, case, dtype, K):
    """UHS_CFG_TEMPORALK for the 2D stars (ring form), the radius-2 box (sum
    form) and the radius-1 box (window form), K = tile_y steps per launch: the result grid is bitwise the
    oracle's for T from 1 to 3K (leading steps, one and several K-step
    launches), halos untouched and every state's halo read from the grid
    that holds it, with edge tiles in both axes, slabs of 3 to 64 rows, 1, 2
    and 4 columns per thread (where compiled) and halos wider than the
    radius."""
    u = gpu
    s = SHAPES[case]
    R = s["radius"]
    st = u.builtin(case).with_coeffs(u.make_coeffs(u.builtin(case), 16))
    dt = np.float64 if dtype == "f64" else np.float32
    TK, NT = u.CFG_TEMPORALK, u.CFG_NT_STORE
    Ts = sorted({1, 2, K - 1, K, K + 1, 2 * K, 2 * K + 1, 3 * K})
    for ext, h in (([67, 301], R), ([23, 45], R + 1), ([150, 40], R)):
        A0, B0 = random_grids(ext, h, 5)
        A0, B0 = A0.astype(dt), B0.astype(dt)
        mask = np.ones(A0.shape, bool)
        mask[tuple(slice(h, n - h) for n in A0.shape)] = False
        for bx, tz, v, fl in ((64, 16, 1, TK), (16, 3, 1, TK | NT), (128, 5, 2, TK), (256, 64, 4, TK | NT),
                              (64, 7, 4, TK), (256, 9, 2, TK)):
            tx = bx - 2 * (K - 1) * R
            if tx < 1 or bx // v < 2 * R:  # the ring needs 2r staging threads
                continue
            cfg = (tx, K, tz, v, fl)
            if v == 4 and case != "2d1r_star" or ("box" in case and v == 2 and dtype == "f64" and R == 2):
                # four columns for the radius-1 star only; boxes 1 or 2 (the
                # FP64 radius-2 box 1)
                with pytest.raises(u.UhsError):
                    u.plan_kind(ext, h, st, {"f64": u.DT_F64, "f32": u.DT_F32, "f32k": u.DT_F32K}[dtype], cfg=cfg)
                continue
            for T in Ts:
                log(case, dtype, ext, h, cfg, "T", T)
                gA, gB = run_gpu(u, st, A0, B0, ext, h, T, dt, dtype == "f32k", cfg=cfg)
                A, B = A0.copy(), B0.copy()
                orc.iterate(A, B, ext, h, s["offsets"], st.coeffs, T, KAHAN if dtype == "f32k" else FUSED)
                got, want = (gA, A) if T % 2 == 0 else (gB, B)
                assert np.array_equal(got, want), (ext, h, cfg, T)
                assert np.array_equal(gA[mask], A0[mask]) and np.array_equal(gB[mask], B0[mask])


def test_k_steps_per_launch_config_checks(gpu):
    u = gpu
    T2, TK, RR, NT = u.CFG_TEMPORAL2, u.CFG_TEMPORALK, u.CFG_ROWRING, u.CFG_NT_STORE
    ext = [64, 96]
    s1, s3 = u.builtin("2d1r_star"), u.builtin("2d3r_star")
    assert u.plan_kind(ext, 1, s1, cfg=(252, 3, 32, 2, TK)) == "win2d"
    assert u.plan_kind(ext, 1, s1, cfg=(242, 8, 32, 2, TK)) == "win2d"
    assert u.plan_kind(ext, 3, s3, cfg=(52, 3, 32, 1, TK | NT)) == "win2d"
    assert u.plan_kind(ext, 1, s1, cfg=(236, 3, 32, 4, TK)) == "win2d"
    for st, e, bad in ((s1, ext, (60, 3, 32, 8, TK)),                   # vec 8
                       (s3, ext, (52, 3, 32, 4, TK)),                   # 4 columns: radius-1 star only
                       (s1, ext, (1020, 3, 32, 1, TK)),                 # 1024 threads
                       (s1, ext, (62, 1, 32, 2, TK)),                   # K = 1
                       (s1, ext, (58, 7, 32, 2, TK)),                   # K = 7 not compiled
                       (s1, ext, (62, 3, 32, 2, TK | T2)),              # with TEMPORAL2
                       (s1, ext, (60, 3, 32, 2, TK | RR)),              # no window form
                       (s1, ext, (61, 3, 32, 2, TK)),                   # tile_x + 4r odd with vec 2
                       (u.builtin("2d1r_box"), ext, (60, 3, 32, 4, TK)),  # box: 1 or 2 columns
                       (u.builtin("2d2r_box"), ext, (60, 3, 32, 2, TK)),  # FP64 radius-2 box: 1 column
                       (u.builtin("2d3r_box"), ext, (52, 3, 32, 1, TK)),
                       (u.builtin("3d1r_star"), [40, 44, 70], (60, 3, 32, 2, TK))):  # 2D only
        with pytest.raises(u.UhsError) as ex:
            u.plan_kind(e, st.radius, st, cfg=bad)
        assert ex.value.code == -3, bad


def test_two_steps_per_launch_config_checks(gpu):
    u = gpu
    st = u.builtin("3d2r_star")
    T2 = u.CFG_TEMPORAL2
    ext = [64, 64, 64]
    assert u.plan_kind(ext, 2, st, cfg=(60, 4, 64, 2, T2)) == "star3d"
    for bad in ((61, 4, 64, 2, T2),                    # tile_x + 2r odd with vec 2
                (64, 8, 64, 4, T2),                    # vec 4
                (124, 8, 64, 2, T2),                   # 64 * 12 threads
                (60, 4, 64, 2, T2 | u.CFG_ROWS2),       # no other variant flags
                (60, 4, 64, 2, T2 | u.CFG_PREFETCH2)):
        with pytest.raises(u.UhsError) as e:
            u.plan_kind(ext, 2, st, cfg=bad)
        assert e.value.code == -3
    with pytest.raises(u.UhsError):
        u.plan_kind(ext, 2, u.builtin("3d2r_box"), cfg=(28, 4, 64, 1, T2))


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("compensated", [False, True])
def test_f32_parity(gpu, orc, case, compensated):
    u = gpu
    s = SHAPES[case]
    st = u.builtin(case).with_coeffs(u.make_coeffs(u.builtin(case), 6))
    ext, h = small_ext(s["dims"]), s["radius"]
    A0, B0 = random_grids(ext, h, 10)
    A0, B0 = A0.astype(np.float32), B0.astype(np.float32)
    gA, gB = run_gpu(u, st, A0, B0, ext, h, 2, np.float32, compensated)
    A, B = A0.copy(), B0.copy()
    orc.iterate(A, B, ext, h, s["offsets"], st.coeffs, 2, KAHAN if compensated else FUSED)
    assert np.array_equal(gA, A) and np.array_equal(gB, B)


def test_general_stencil_paths(gpu, orc):
    """Stencils outside the 12 built-in forms: up to 32 terms run the
    streaming program kernel, larger ones the general one-thread-per-point
    kernel; FP64, FP32 and FP32 compensated all bitwise equal to the
    oracle."""
    u = gpu
    rng = np.random.default_rng(3)
    cases = [
        # the 7-point star in another term order
        (3, [[0, 0, 0], [0, 0, 1], [0, 0, -1], [0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0]]),
        # one-sided ("arbitrary shape") 3D stencil
        (3, [[0, 0, 0], [-1, 0, 0], [0, -2, 0], [0, 0, 1], [1, 1, 1], [-2, 0, 2]]),
        # 2D upwind-like stencil
        (2, [[0, 0], [0, -1], [-1, 0], [-2, 0], [0, -3]]),
        # 40 distinct offsets within radius 3: more terms than the program kernel takes
        (3, [list(o) for o in sorted({tuple(int(v) for v in rng.integers(-3, 4, 3)) for _ in range(200)})[:40]]),
    ]
    for dims, offs in cases:
        c = rng.random(len(offs))
        st = u.StencilSpec(dims, offs, c)
        h = 3
        ext = [23, 31, 45] if dims == 3 else [50, 77]
        kind = "program" if len(offs) <= 32 else "gather"
        assert u.plan_kind(ext, h, st) == kind and u.plan_kind(ext, h, st, u.DT_F32) == kind
        log("general stencil", dims, "D", len(offs), "terms")
        A0, B0 = random_grids(ext, h, 4)
        gA, gB = run_gpu(u, st, A0, B0, ext, h, 3)
        A, B = A0.copy(), B0.copy()
        orc.iterate(A, B, ext, h, offs, c, 3, FUSED)
        assert np.array_equal(gA, A) and np.array_equal(gB, B)
        for compensated, mode in ((False, FUSED), (True, KAHAN)):
            gA, gB = run_gpu(u, st, A0, B0, ext, h, 2, dtype=np.float32, compensated=compensated)
            A, B = A0.astype(np.float32), B0.astype(np.float32)
            orc.iterate(A, B, ext, h, offs, c, 2, mode)
            assert np.array_equal(gA, A) and np.array_equal(gB, B)


def test_config1_3d7pt_256_T10(gpu, orc):
    """3d1r_star, 256^3, c_center=0.4, c_face=0.1, U[0,1) interior, zero
    halo, T=10, against the oracle."""
    u = gpu
    s = SHAPES["3d1r_star"]
    c = [0.4 if not any(o) else 0.1 for o in s["offsets"]]
    st = u.builtin("3d1r_star").with_coeffs(c)
    ext, h = [256, 256, 256], 1
    A0 = orc.fill(ext, h, 1)
    B0 = np.zeros_like(A0)
    gA, gB = run_gpu(u, st, A0, B0, ext, h, 10)
    A, B = A0.copy(), B0.copy()
    orc.iterate(A, B, ext, h, s["offsets"], c, 10, FUSED)
    assert np.array_equal(gA, A)
    A, B = A0.copy(), B0.copy()
    orc.iterate(A, B, ext, h, s["offsets"], c, 10, PLAIN)
    err = rel_err(gA, A)
    log("config1 max rel err vs plain oracle", err)
    assert err <= 1e-12


def test_error_norms(gpu, orc):
    u = gpu
    ext, h = [30, 40, 50], 2
    a = orc.fill(ext, h, 1)
    b = orc.fill(ext, h, 2)
    da, db = u.DeviceArray.from_host(a), u.DeviceArray.from_host(b)
    a32 = u.DeviceArray.from_host(a.astype(np.float32))
    m, d2, r2 = u.error_norms(da, db, ext, h)
    ia = a[h:-h, h:-h, h:-h]
    ib = b[h:-h, h:-h, h:-h]
    assert abs(m - np.abs(ia - ib).max()) == 0
    assert abs(d2 - np.sqrt(((ia - ib) ** 2).sum())) <= 1e-12 * d2
    assert abs(r2 - np.sqrt((ib ** 2).sum())) <= 1e-12 * r2
    m32, _, _ = u.error_norms(a32, db, ext, h)
    assert m32 == np.abs(ia.astype(np.float32).astype(np.float64) - ib).max()
    for d in (da, db, a32):
        d.close()


def test_fma_probe_exact(gpu):
    # x = 0.5, y = 0.25, a_i starting at i: three rounds stay exact in
    # binary64, so every thread's sum is known in closed form:
    # a_i -> i/8 + 0.25/4 + 0.25/2 + 0.25.
    u = gpu
    n = 1000
    out = u.DeviceArray([n], np.float64)
    u.check(u.lib().uhs_fma_f64(out.ptr, n, 0.5, 0.25, 3, None))
    got = out.download()
    out.close()
    want = sum(i / 8 + 0.0625 + 0.125 + 0.25 for i in range(32))
    assert np.all(got == want)


def test_argument_errors(gpu):
    u = gpu
    st = u.builtin("3d1r_star")
    A = u.DeviceArray([10, 10, 10], np.float64)
    with pytest.raises(u.UhsError) as e:
        u.sweep_f64(A, A, [8, 8, 8], 1, st, 1)
    assert e.value.code == -1
    with pytest.raises(u.UhsError) as e:
        u.sweep_f64(A, A, [8, 8, 8], 1, st, 1, cfg=(64, 8, 64, 3))
    assert e.value.code == -3
    A.close()
