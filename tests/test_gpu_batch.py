"""The batched path (one hipGraph of persistent work-queue launches, star
grids temporally blocked) against the oracle: each grid's result buffer is
bitwise equal to T single sweeps of the oracle's fused mode, for uneven
extents, mixed stencils, nonzero halos that differ between A and B, every
temporal depth and odd/even T."""
import sys

import numpy as np
import pytest

from tests.oracle_py import FUSED

pytestmark = pytest.mark.gpu


def make_grids(u, orc, kinds, seed, halo_extra=0):
    rng = np.random.default_rng(seed)
    out = []
    for i, case in enumerate(kinds):
        st = u.builtin(case)
        st = st.with_coeffs(u.make_coeffs(st, 100 + i))
        h = st.radius + halo_extra
        ext = [int(rng.integers(5, 41)) for _ in range(3)]
        shape = [e + 2 * h for e in ext]
        A0 = rng.random(shape)
        B0 = rng.random(shape)
        out.append((case, st, ext, h, A0, B0))
    return out


def check_batch(u, orc, grids, T, cfg):
    dev = []
    for case, st, ext, h, A0, B0 in grids:
        dev.append((u.DeviceArray.from_host(A0), u.DeviceArray.from_host(B0), ext, h, st))
    b = u.Batch(dev, T, cfg)
    print("batch", cfg, b.info(), file=sys.stderr, flush=True)
    b.run()
    u.synchronize()
    for (case, st, ext, h, A0, B0), (dA, dB, _, _, _) in zip(grids, dev):
        A, B = A0.copy(), B0.copy()
        orc.iterate(A, B, ext, h, st.offsets, st.coeffs, T, FUSED)
        res, ref = (dA.download(), A) if T % 2 == 0 else (dB.download(), B)
        assert np.array_equal(res, ref), (case, ext, h, T, cfg)
    b.close()
    for dA, dB, _, _, _ in dev:
        dA.close()
        dB.close()


# Box tile forms: 256 points one row per thread, 512 points two rows per
# thread, each with one or two planes of loads in flight.
BOX_FORMS = {1: (32, 8, 9, 0), 2: (32, 16, 9, 1), 5: (32, 8, 9, 2), 8: (16, 32, 5, 3)}


@pytest.mark.parametrize("depth", [1, 2, 4])
@pytest.mark.parametrize("T", [1, 2, 5, 8])
def test_batch_mix_star_box(gpu, orc, depth, T):
    """Star tile 16x8: register-queue two-step tiles at depth 2 (its ghost-
    zone extension needs fewer threads than the workgroup has)."""
    u = gpu
    kinds = ["3d1r_star", "3d2r_box"] * 4
    grids = make_grids(u, orc, kinds, seed=depth * 10 + T)
    cfg = u.batch_default_config()
    cfg.star_depth = depth
    cfg.star_tile_x, cfg.star_tile_y, cfg.star_tile_z = 16, 8, 7
    cfg.box_tile_x, cfg.box_tile_y, cfg.box_tile_z, cfg.flags = BOX_FORMS[T]
    check_batch(u, orc, grids, T, cfg)


@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("star_tile", [(64, 8), (62, 6), (30, 14)])
def test_batch_star_tile_forms(gpu, orc, depth, star_tile):
    """64x8: register-queue single steps, LDS-ring two-step tiles; 62x6 and
    30x14: register-queue two-step tiles that fill the workgroup exactly."""
    u = gpu
    kinds = ["3d1r_star", "3d2r_box", "3d1r_star"]
    grids = make_grids(u, orc, kinds, seed=31 + depth)
    cfg = u.batch_default_config()
    cfg.star_depth = depth
    cfg.star_tile_x, cfg.star_tile_y, cfg.star_tile_z = star_tile + (11,)
    cfg.flags = u.BATCH_BOX_ROWS2 | u.BATCH_BOX_PLANES2  # shared launches
    check_batch(u, orc, grids, 7, cfg)


@pytest.mark.parametrize("depth", [1, 2, 4])
def test_batch_all_3d_kinds_wide_halo(gpu, orc, depth):
    u = gpu
    kinds = ["3d1r_star", "3d2r_star", "3d3r_star", "3d1r_box", "3d2r_box", "3d3r_box"]
    grids = make_grids(u, orc, kinds, seed=7 + depth, halo_extra=1)
    cfg = u.batch_default_config()
    cfg.star_depth = depth
    cfg.star_tile_x, cfg.star_tile_y, cfg.star_tile_z = 32, 4, 16
    cfg.flags = u.BATCH_BOX_ROWS2 | u.BATCH_BOX_PLANES2  # shared launches
    check_batch(u, orc, grids, 6, cfg)


def test_batch_default_config_and_rejections(gpu, orc):
    u = gpu
    d = u.batch_default_config()
    assert d.flags & u.BATCH_SPLIT and d.flags & u.BATCH_CONCURRENT
    grids = make_grids(u, orc, ["3d1r_star", "3d2r_box", "3d1r_star"], seed=3)
    check_batch(u, orc, grids, 4, None)
    # other stencil sets: the shared-launch default
    check_batch(u, orc, make_grids(u, orc, ["3d1r_star", "3d1r_box", "3d2r_box"], seed=4), 4, None)
    case, st, ext, h, A0, B0 = grids[0]
    dA = u.DeviceArray.from_host(A0)
    dB = u.DeviceArray.from_host(B0)
    bad = u.batch_default_config()
    bad.star_depth = 3
    with pytest.raises(u.UhsError) as e:
        u.Batch([(dA, dB, ext, h, st)], 2, bad)
    assert e.value.code == -3
    # flags 16: no such flag; 8: UHS_BATCH_CONCURRENT without UHS_BATCH_SPLIT
    for tile in ((32, 8, 9, 1), (32, 15, 9, 1), (32, 16, 9, 0), (32, 8, 9, 16), (32, 8, 9, 8)):
        bad = u.batch_default_config()
        bad.box_tile_x, bad.box_tile_y, bad.box_tile_z, bad.flags = tile
        with pytest.raises(u.UhsError) as e:
            u.Batch([(dA, dB, ext, h, st)], 2, bad)
        assert e.value.code == -3
    two_d = u.builtin("2d1r_star")
    with pytest.raises(u.UhsError):
        u.Batch([(dA, dB, ext, h, two_d)], 2)
    dA.close()
    dB.close()


@pytest.mark.parametrize("T", [3, 4, 5])
@pytest.mark.parametrize("depth", [1, 2, 4])
def test_batch_rerun_continues(gpu, orc, T, depth):
    """Runs of one batch continue from the state the last run left: three
    runs equal 3T single sweeps, also for odd T (whose odd-numbered runs
    start from B, through the second graph with the roles swapped)."""
    u = gpu
    grids = make_grids(u, orc, ["3d1r_star", "3d2r_box", "3d2r_star"], seed=50 + T + depth, halo_extra=1)
    dev = [(u.DeviceArray.from_host(A0), u.DeviceArray.from_host(B0), ext, h, st)
           for case, st, ext, h, A0, B0 in grids]
    cfg = u.batch_default_config()
    cfg.star_depth = depth
    cfg.star_tile_x, cfg.star_tile_y, cfg.star_tile_z = 16, 8, 7
    b = u.Batch(dev, T, cfg)
    runs = 3
    for _ in range(runs):
        b.run()
    u.synchronize()
    b.close()
    for (case, st, ext, h, A0, B0), (dA, dB, _, _, _) in zip(grids, dev):
        A, B = A0.copy(), B0.copy()
        orc.iterate(A, B, ext, h, st.offsets, st.coeffs, runs * T, FUSED)
        res, ref = (dA.download(), A) if (runs * T) % 2 == 0 else (dB.download(), B)
        assert np.array_equal(res, ref), (case, ext, h, T, depth)
        dA.close()
        dB.close()


@pytest.mark.parametrize("mode", ["split", "concurrent"])
@pytest.mark.parametrize("depth,star_tile", [(1, (32, 16)), (1, (64, 16)), (2, (62, 14)), (2, (30, 14)),
                                             (2, (32, 12)), (4, (32, 8))])
def test_batch_split(gpu, orc, mode, depth, star_tile):
    """Star grids in launches of their own (512-thread workgroups for the
    benchmark mix), in one chain with the box launches or as a concurrent
    second chain: bitwise the oracle's T single sweeps, odd and even T."""
    u = gpu
    kinds = ["3d1r_star", "3d2r_box"] * 3
    for T in (5, 6):
        grids = make_grids(u, orc, kinds, seed=70 + depth + T)
        cfg = u.batch_default_config()
        cfg.star_depth = depth
        cfg.star_tile_x, cfg.star_tile_y, cfg.star_tile_z = star_tile + (9,)
        cfg.flags = u.BATCH_BOX_ROWS2 | u.BATCH_BOX_PLANES2 | u.BATCH_SPLIT
        if mode == "concurrent":
            cfg.flags |= u.BATCH_CONCURRENT
        check_batch(u, orc, grids, T, cfg)


def test_batch_split_other_stencils(gpu, orc):
    """Split batches of other stencil sets keep 256-thread star launches."""
    u = gpu
    kinds = ["3d1r_star", "3d2r_star", "3d1r_box", "3d2r_box"]
    grids = make_grids(u, orc, kinds, seed=91, halo_extra=1)
    cfg = u.batch_default_config()
    cfg.star_tile_x, cfg.star_tile_y, cfg.star_tile_z = 16, 8, 7
    cfg.flags = u.BATCH_BOX_ROWS2 | u.BATCH_SPLIT | u.BATCH_CONCURRENT
    check_batch(u, orc, grids, 5, cfg)
