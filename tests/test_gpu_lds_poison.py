"""The LDS-poison debug mode (uhs_set_lds_poison): every tile fills the LDS
span it uses with a NaN pattern before staging, so a read of a cell a kernel
never wrote would turn into NaN in its output.  Under the mode, every kernel
that stages through LDS must still give results bitwise equal to the
oracle's fused mode: the one-step kernels of all 12 stencils (partial tiles,
random halos), the two-step 3D star, 3D box and 2D tiles (interior and edge paths), the box
tile forms, the FP32 forms, the batched work queue (register and LDS-ring
star tiles, box tiles) and the fused-program streaming kernel."""
import json
import os
import sys

import numpy as np
import pytest

from tests.oracle_py import FUSED, KAHAN
from tests.test_gpu_batch import check_batch, make_grids
from tests.test_program import random_program

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = json.load(open(os.path.join(HERE, "golden", "shapes.json")))
T2, NT, PF2, ROWS2, PL2, RR = 32, 1, 2, 16, 64, 128


def log(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture
def poisoned(gpu):
    units = gpu.set_lds_poison(True)
    # One flag per kernel source file (sweeps, two-step, boxes, 2D, batch,
    # programs, ...).
    assert units >= 9
    yield gpu
    gpu.set_lds_poison(False)


def sweep(u, st, A0, B0, ext, h, T, dtype=np.float64, compensated=False, cfg=None):
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


def check(u, orc, case, ext, T, cfg=None, dtype=np.float64, mode=FUSED, seed=5):
    s = SHAPES[case]
    st = u.builtin(case)
    c = u.make_coeffs(st, 77)
    st = st.with_coeffs(c)
    h = s["radius"]
    rng = np.random.default_rng(seed)
    shape = [e + 2 * h for e in ext]
    A0, B0 = rng.random(shape).astype(dtype), rng.random(shape).astype(dtype)
    log("poison", case, ext, "T", T, "cfg", cfg, np.dtype(dtype).name, u.plan_kind(ext, h, st, cfg=cfg))
    gA, gB = sweep(u, st, A0, B0, ext, h, T, dtype, mode == KAHAN, cfg)
    A, B = A0.copy(), B0.copy()
    orc.iterate(A, B, ext, h, s["offsets"], c, T, mode)
    # The result grid (A for even T, B for odd T) bitwise, and the halos of
    # both untouched.  With two steps per launch the other grid's interior
    # holds an earlier state than the one-step ping-pong leaves there.
    got, want = (gA, A) if T % 2 == 0 else (gB, B)
    assert np.array_equal(got, want), (case, ext, T, cfg)
    mask = np.ones(A0.shape, bool)
    mask[tuple(slice(h, n - h) for n in A0.shape)] = False
    assert np.array_equal(gA[mask], A0[mask]) and np.array_equal(gB[mask], B0[mask])


@pytest.mark.parametrize("case", list(SHAPES))
def test_default_kernels_under_poison(poisoned, orc, case):
    ext = [67, 301] if SHAPES[case]["dims"] == 2 else [19, 37, 70]
    two_step = (case.startswith("3d") and "star" in case) or (case.startswith("2d") and case != "2d3r_box") or \
        case == "3d1r_box"
    for T in ((1, 3, 4) if two_step else (1, 3)):
        check(poisoned, orc, case, ext, T)


@pytest.mark.parametrize("case", ["3d1r_star", "3d2r_star", "3d3r_star", "3d1r_box"])
def test_two_step_tiles_under_poison(poisoned, orc, case):
    # Grids wide enough for interior tiles next to edge tiles.
    for cfg in ((10, 4, 5, 2, T2), (30, 6, 7, 1, T2 | NT), (14, 6, 3, 2, T2)):
        check(poisoned, orc, case, [40, 44, 70], 5, cfg)


@pytest.mark.parametrize("case", ["2d1r_star", "2d2r_star", "2d3r_star", "2d1r_box", "2d2r_box"])
@pytest.mark.parametrize("K", [3, 4, 8])
def test_k_step_tiles_2d_under_poison(poisoned, orc, case, K):
    R = SHAPES[case]["radius"]
    TK = poisoned.CFG_TEMPORALK
    for bx, tz, v, fl in ((64, 16, 1, TK), (16, 5, 1, TK | NT), (128, 7, 2, TK), (256, 9, 4, TK)):
        tx = bx - 2 * (K - 1) * R
        if tx < 1 or bx // v < 2 * R or ("box" in case and v > 1) or (v == 4 and R > 1):
            # FP64 boxes here one column; four columns for the radius-1 star only
            continue
        check(poisoned, orc, case, [150, 301], 2 * K + 1, (tx, K, tz, v, fl))


@pytest.mark.parametrize("case", ["2d1r_star", "2d2r_star", "2d3r_star", "2d1r_box", "2d2r_box", "2d3r_box"])
def test_two_step_tiles_2d_under_poison(poisoned, orc, case):
    R = SHAPES[case]["radius"]
    if case == "2d3r_box":  # the sum form only (no FP64 window form at radius 3)
        for bx, tz, v, fl in ((64, 16, 1, T2), (16, 5, 1, T2 | NT), (128, 7, 2, T2)):
            check(poisoned, orc, case, [150, 301], 5, (bx - 2 * R, 1, tz, v, fl))
        return
    # without UHS_CFG_ROWRING the ring form (stars) or the sum form (boxes);
    # with it the window form
    for form in (0, RR):
        for bx, tz, v, fl in ((64, 16, 1, T2), (16, 5, 1, T2 | NT), (128, 7, 2 if R == 1 or not form else 1, T2)):
            check(poisoned, orc, case, [150, 301], 5, (bx - 2 * R, 1, tz, v, fl | form))


@pytest.mark.parametrize("case", ["3d1r_box", "3d2r_box", "3d3r_box", "3d1r_star", "3d2r_star"])
def test_tile_forms_under_poison(poisoned, orc, case):
    box = "box" in case
    cfgs = ([(32, 8, 9, 1, 0), (32, 16, 9, 1, ROWS2), (32, 16, 9, 1, ROWS2 | PL2), (64, 4, 9, 2, PL2)] if box
            else [(64, 8, 9, 4, 0), (128, 4, 9, 8, PF2 | NT), (32, 8, 9, 1, 8)])
    for cfg in cfgs:
        try:
            poisoned.plan_kind([23, 41, 70], SHAPES[case]["radius"], poisoned.builtin(case), cfg=cfg)
        except poisoned.UhsError:
            continue
        check(poisoned, orc, case, [23, 41, 70], 2, cfg)


@pytest.mark.parametrize("case", ["3d2r_star", "3d3r_box", "2d3r_box", "3d1r_star"])
def test_fp32_forms_under_poison(poisoned, orc, case):
    ext = [67, 301] if case.startswith("2d") else [19, 37, 70]
    check(poisoned, orc, case, ext, 2, dtype=np.float32, mode=FUSED)
    check(poisoned, orc, case, ext, 2, dtype=np.float32, mode=KAHAN)


@pytest.mark.parametrize("depth", [1, 2, 4])
def test_batch_under_poison(poisoned, orc, depth):
    u = poisoned
    grids = make_grids(u, orc, ["3d1r_star", "3d2r_box", "3d1r_star", "3d2r_box"], seed=200 + depth)
    cfg = u.batch_default_config()
    cfg.star_depth = depth
    cfg.star_tile_x, cfg.star_tile_y, cfg.star_tile_z = 16, 8, 7
    check_batch(u, orc, grids, 5, cfg)
    # The default configuration of the mix: split, concurrent chains.
    check_batch(u, orc, make_grids(u, orc, ["3d1r_star", "3d2r_box"] * 2, seed=300 + depth), 4, None)


@pytest.mark.parametrize("seed", range(4))
def test_program_stream_under_poison(poisoned, orc, seed):
    u = poisoned
    rng = np.random.default_rng(500 + seed)
    dims = 2 + seed % 2
    shape, prods, offs, coefs, srcs, lo, hi = random_program(rng, dims)
    prog = u.ProgramSpec(dims, prods, offs, coefs, srcs, lo, hi)
    A0, B0 = rng.random(shape), rng.random(shape)
    log("poison program", seed, shape, len(prods))
    A = u.DeviceArray.from_host(A0)
    B = u.DeviceArray.from_host(B0)
    try:
        u.program_sweep_f64(A, B, list(shape), prog, 3, kernel=1)
        u.synchronize()
        gA, gB = A.download(), B.download()
    except u.UhsError as e:
        assert e.code == -3, e  # frame beyond the LDS: refused before any launch
        return
    finally:
        A.close()
        B.close()
    Ar, Br = A0.copy(), B0.copy()
    assert orc.program(Ar, Br, shape, prods, offs, coefs, srcs, lo, hi, 3, FUSED) == 0
    assert np.array_equal(gA, Ar) and np.array_equal(gB, Br)


def test_poison_off_again(gpu):
    assert gpu.set_lds_poison(False) >= 9
