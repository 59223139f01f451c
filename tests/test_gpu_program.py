"""Fused two-level programs on the GPU (uhs_program_sweep_f64, kernel
k_program) against the oracle's nested evaluation (orc_program_f64).

Bar: bitwise equality with the oracle's fused mode on both grids (so cells
outside the domain are checked to be untouched), T = 1, 2, 3 ping-pong; and
within 1e-12 of the terms' magnitudes against the plain mode, the reference's
separate multiply and add."""
import sys

import numpy as np
import pytest

from tests.oracle_py import FUSED, KAHAN, PLAIN
from tests.test_program import NESTED_2D, ONE_SIDED_2D, max_program, random_program, run_parsed

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(u, prog, A0, B0, T, kernel=0, tile=None):
    A = u.DeviceArray.from_host(A0)
    B = u.DeviceArray.from_host(B0)
    try:
        u.program_sweep_f64(A, B, list(A0.shape), prog, T, kernel=kernel, tile=tile)
        u.synchronize()
        return A.download(), B.download()
    finally:
        A.close()
        B.close()


# (kernel, 2D tile, 3D tile): the library's choice, the pointwise kernel,
# and streamed tiles that leave partial tiles and slabs at every edge (one
# cell per tile too; rows of two 64-cell chunks).
KERNELS = [(0, None), (2, None), (1, (60, 16, 1), (60, 8, 4)), (1, (7, 3, 1), (5, 3, 2)), (1, (1, 1, 1), (1, 1, 1)),
           (1, (200, 5, 1), (100, 2, 7))]


def kernel_tile(kt, dims):
    if kt[0] != 1:
        return kt[0], None
    return 1, list(kt[1] if dims == 2 else kt[2])


@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("kt", KERNELS, ids=["auto", "pointwise", "tiled", "tiled_small", "tiled_1", "tiled_wide"])
def test_random_programs_bitwise(gpu, orc, seed, kt):
    u = gpu
    rng = np.random.default_rng(seed)
    dims = 2 + seed % 2
    shape, prods, offs, coefs, srcs, lo, hi = random_program(rng, dims)
    prog = u.ProgramSpec(dims, prods, offs, coefs, srcs, lo, hi)
    kernel, tile = kernel_tile(kt, dims)
    A0, B0 = rng.random(shape), rng.random(shape)
    log("program seed", seed, "shape", shape, "producers", len(prods), "terms", len(offs), "kernel", kernel, tile)
    for T in (1, 2, 3):
        try:
            gA, gB = run_gpu(u, prog, A0, B0, T, kernel, tile)
        except u.UhsError as e:
            # A frame beyond the LDS is refused before any launch.
            assert e.code == -3 and kernel == 1, e
            return
        A, B = A0.copy(), B0.copy()
        assert orc.program(A, B, shape, prods, offs, coefs, srcs, lo, hi, T, FUSED) == 0
        assert np.array_equal(gA, A) and np.array_equal(gB, B), (seed, T)


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("kernel", [0, 1, 2])
def test_max_program_bitwise(gpu, orc, dims, kernel):
    """4 producers of 32 terms, 32 output terms (both producer tables of the
    tiled kernel in use)."""
    u = gpu
    rng = np.random.default_rng(40 + dims)
    shape, prods, offs, coefs, srcs, lo, hi = max_program(rng, dims)
    prog = u.ProgramSpec(dims, prods, offs, coefs, srcs, lo, hi)
    A0, B0 = rng.random(shape), rng.random(shape)
    log("max program", dims, "D, kernel", kernel)
    for T in (1, 2):
        gA, gB = run_gpu(u, prog, A0, B0, T, kernel)
        A, B = A0.copy(), B0.copy()
        assert orc.program(A, B, shape, prods, offs, coefs, srcs, lo, hi, T, FUSED) == 0
        assert np.array_equal(gA, A) and np.array_equal(gB, B), T


@pytest.mark.parametrize("kernel", [1, 2])
def test_nested_2d_against_plain(gpu, orc, kernel):
    u = gpu
    prog, info = u.parse_dsl_program(NESTED_2D)
    rng = np.random.default_rng(21)
    A0 = rng.random((98, 98))
    log("nested 2d program, T =", info.iteration, "kernel", kernel)
    gA, gB = run_gpu(u, prog, A0, A0.copy(), info.iteration, kernel)
    A, B = A0.copy(), A0.copy()
    assert run_parsed(orc, prog, A, B, info.iteration, FUSED) == 0
    assert np.array_equal(gA, A) and np.array_equal(gB, B)
    A, B = A0.copy(), A0.copy()
    assert run_parsed(orc, prog, A, B, info.iteration, PLAIN) == 0
    # The coefficients' magnitudes sum to 1.7 and 1.3: values stay O(1).
    assert np.max(np.abs(gA - A)) <= 1e-12


@pytest.mark.parametrize("kernel", [0, 2])
def test_large_3d_grid(gpu, orc, kernel):
    """27M written cells: 16,875 tiled blocks; for the pointwise kernel more
    than its 65,536 x 256 threads, so every thread takes several cells."""
    u = gpu
    shape, lo, hi = [306, 306, 306], [3, 3, 3], [303, 303, 303]
    prods = [([[-1, 0, 0], [0, -1, 0], [0, 0, -1], [0, 0, 0], [0, 0, 1], [0, 1, 0], [1, 0, 0]],
              [0.5, 0.2, 0.4, 0.5, 0.6, 0.8, 0.5])]
    offs = [[-1, -1, -1], [0, 0, 0], [1, 1, 1], [2, -2, 0]]
    coefs, srcs = [0.1, 0.5, 0.9, -0.25], [0, 0, 0, -1]
    prog = u.ProgramSpec(3, prods, offs, coefs, srcs, lo, hi)
    rng = np.random.default_rng(8)
    A0 = rng.random(shape)
    B0 = rng.random(shape)
    log("program 3d", shape, "T = 1, kernel", kernel)
    gA, gB = run_gpu(u, prog, A0, B0, 1, kernel)
    A, B = A0.copy(), B0.copy()
    assert orc.program(A, B, shape, prods, offs, coefs, srcs, lo, hi, 1, FUSED) == 0
    assert np.array_equal(gB, B) and np.array_equal(gA, A0)


def test_far_reach_falls_back_to_pointwise(gpu, orc):
    """Terms 40 planes and 8 rows away: no tiled frame fits in the LDS, so
    the library takes the pointwise kernel."""
    u = gpu
    shape, lo, hi = [90, 20, 70], [41, 9, 2], [49, 11, 68]
    prods = [([[0, 0, 1], [0, 0, -1]], [0.5, 0.25])]
    offs = [[-40, 0, 0], [40, 0, 0], [0, -8, 0], [0, 8, 0]]
    coefs, srcs = [0.3, 0.6, -0.2, 0.1], [0, -1, 0, -1]
    prog = u.ProgramSpec(3, prods, offs, coefs, srcs, lo, hi)
    rng = np.random.default_rng(4)
    A0, B0 = rng.random(shape), rng.random(shape)
    gA, gB = run_gpu(u, prog, A0, B0, 2)
    A, B = A0.copy(), B0.copy()
    assert orc.program(A, B, shape, prods, offs, coefs, srcs, lo, hi, 2, FUSED) == 0
    assert np.array_equal(gA, A) and np.array_equal(gB, B)


@pytest.mark.parametrize("kernel", [1, 2])
def test_one_sided_programs(gpu, orc, kernel):
    """Domains that touch the grid's edge on the side a program does not
    read (the reference's arbitrary-shape examples, at small sizes)."""
    u = gpu
    prog, info = u.parse_dsl_program(ONE_SIDED_2D)
    rng = np.random.default_rng(31)
    A0, B0 = rng.random((66, 66)), np.zeros((66, 66))
    gA, gB = run_gpu(u, prog, A0, B0, 4, kernel)
    A, B = A0.copy(), B0.copy()
    assert run_parsed(orc, prog, A, B, 4, FUSED) == 0
    assert np.array_equal(gA, A) and np.array_equal(gB, B)
    offs = [[2, 0, 0], [1, 0, 0], [0, 0, 0], [0, 0, 1], [0, 0, 2], [0, 1, 0], [0, 2, 0]]
    coefs = [0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7]
    shape, lo, hi = [34, 36, 70], [0, 0, 0], [32, 34, 68]
    prog3 = u.ProgramSpec(3, [], offs, coefs, [-1] * 7, lo, hi)
    A0, B0 = rng.random(shape), np.zeros(shape)
    log("one-sided 3d, kernel", kernel)
    gA, gB = run_gpu(u, prog3, A0, B0, 2, kernel)
    A, B = A0.copy(), B0.copy()
    assert orc.program(A, B, shape, [], offs, coefs, [-1] * 7, lo, hi, 2, FUSED) == 0
    assert np.array_equal(gA, A) and np.array_equal(gB, B)


def chain_program(dims):
    """A 4-deep chain: producer 0 reads the input, 1 reads 0 and the input,
    2 reads 1 and 0, 3 reads 2; the output reads 3, 1 and the input."""
    def o(*v):
        return list(v) if dims == 3 else list(v[1:])
    prods = [
        ([o(0, 0, 1), o(0, -1, 0), o(1, 0, 0)], [0.5, 0.25, -0.125], [-1, -1, -1]),
        ([o(0, 0, 0), o(0, 0, -1), o(-1, 1, 0)], [0.75, 0.5, 0.25], [0, -1, 0]),
        ([o(0, 1, 0), o(0, 0, 1), o(1, 0, -1)], [-0.5, 1.5, 0.25], [1, 0, 1]),
        ([o(0, 0, -1), o(-1, 0, 0)], [0.5, 0.5], [2, 2]),
    ]
    offs = [o(0, 0, 0), o(0, 1, -1), o(1, 0, 1), o(0, 0, 2)]
    shape = [18, 20, 90] if dims == 3 else [40, 150]
    return shape, prods, offs, [0.5, -0.25, 0.125, 1.0], [3, 1, 3, -1], [5] * dims, [n - 5 for n in shape]


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("kernel", [0, 1, 2])
def test_chain_program_bitwise(gpu, orc, dims, kernel):
    u = gpu
    shape, prods, offs, coefs, srcs, lo, hi = chain_program(dims)
    prog = u.ProgramSpec(dims, prods, offs, coefs, srcs, lo, hi)
    rng = np.random.default_rng(77 + dims)
    A0, B0 = rng.random(shape), rng.random(shape)
    log("chain program", dims, "D, kernel", kernel)
    for T in (1, 2, 3):
        gA, gB = run_gpu(u, prog, A0, B0, T, kernel)
        A, B = A0.copy(), B0.copy()
        assert orc.program(A, B, shape, prods, offs, coefs, srcs, lo, hi, T, FUSED) == 0
        assert np.array_equal(gA, A) and np.array_equal(gB, B), T


def test_tuned_tile_bitwise(gpu, orc):
    """The program tuner picks a streaming tile that fits; that tile's
    result is the oracle's."""
    u = gpu
    prog, info = u.parse_dsl_program(NESTED_2D)
    rng = np.random.default_rng(13)
    A0 = rng.random((98, 98))
    A = u.DeviceArray.from_host(A0)
    B = u.DeviceArray.from_host(A0)
    try:
        tile, ms, n = u.tune_program_f64(A, B, [98, 98], prog, reps=2)
    finally:
        A.close()
        B.close()
    log("tuned tile", tile, ms, n)
    assert n > 0 and ms > 0 and tile[2] == 1
    gA, gB = run_gpu(u, prog, A0, A0.copy(), 4, 1, tile)
    # The tuned tile is now the untiled call's choice; same result.
    gA2, gB2 = run_gpu(u, prog, A0, A0.copy(), 4)
    assert np.array_equal(gA, gA2) and np.array_equal(gB, gB2)
    A, B = A0.copy(), A0.copy()
    assert run_parsed(orc, prog, A, B, 4, FUSED) == 0
    assert np.array_equal(gA, A) and np.array_equal(gB, B)
    # Nothing fits: far-reaching terms leave only the pointwise kernel.
    far = u.ProgramSpec(3, [], [[-40, 0, 0], [40, 0, 0], [0, -8, 0], [0, 8, 0]], [0.5, 0.5, 0.1, 0.1],
                        [-1] * 4, [41, 9, 1], [49, 11, 69])
    A = u.DeviceArray([90, 20, 70], np.float64)
    B = u.DeviceArray([90, 20, 70], np.float64)
    try:
        with pytest.raises(u.UhsError, match="no tile"):
            u.tune_program_f64(A, B, [90, 20, 70], far)
    finally:
        A.close()
        B.close()


def run_gpu_f32(u, prog, A0, B0, T, compensated, kernel=0, tile=None):
    A = u.DeviceArray.from_host(A0.astype(np.float32))
    B = u.DeviceArray.from_host(B0.astype(np.float32))
    try:
        u.program_sweep_f32(A, B, list(A0.shape), prog, T, compensated, kernel=kernel, tile=tile)
        u.synchronize()
        return A.download(), B.download()
    finally:
        A.close()
        B.close()


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("kernel", [0, 2])
@pytest.mark.parametrize("compensated", [False, True])
def test_f32_programs_bitwise(gpu, orc, seed, kernel, compensated):
    """FP32 storage and arithmetic (plain fused sums, or Kahan-compensated),
    chains included, against the oracle's FP32 modes."""
    u = gpu
    rng = np.random.default_rng(300 + seed)
    dims = 2 + seed % 2
    shape, prods, offs, coefs, srcs, lo, hi = random_program(rng, dims)
    prog = u.ProgramSpec(dims, prods, offs, coefs, srcs, lo, hi)
    A0, B0 = rng.random(shape).astype(np.float32), rng.random(shape).astype(np.float32)
    log("f32 program seed", seed, "kernel", kernel, "compensated", compensated)
    for T in (1, 2):
        try:
            gA, gB = run_gpu_f32(u, prog, A0, B0, T, compensated, kernel)
        except u.UhsError as e:
            assert e.code == -3 and kernel == 1, e
            return
        A, B = A0.copy(), B0.copy()
        assert orc.program(A, B, shape, prods, offs, coefs, srcs, lo, hi, T, KAHAN if compensated else FUSED) == 0
        assert np.array_equal(gA, A) and np.array_equal(gB, B), (seed, T)
