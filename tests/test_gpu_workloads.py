"""The benchmark workloads themselves, at their own sizes, through the C ABI.

* The headline sweep (3d1r_star FP64 512^3, the library's default: two
  steps per launch), the radius-2 star's, the radius-1 box's and the 2D
  stencils' two-step defaults (8192^2), bitwise equal to the oracle's fused mode at T = 4 (two
  two-step launches), and at T = 100 bitwise equal to one launch per step;
  the FP32 two-step defaults (plain and compensated) the same at T = 4.
* All 12 stencils at 512^3 / 8192^2 and T = 4 with dyadic symmetric
  coefficients on a linear field: the field is a fixed point of the sweep
  (every product and partial sum is exact, and a symmetric stencil whose
  coefficients sum to 1 maps a linear function to itself), so every cell of
  both grids must come back bit for bit -- for the star defaults this runs
  the two-step kernel at full size, whose ghost-zone levels take the halo.
* The uneven batch exactly as the benchmark builds it (64 grids, extents
  64-384, half 7-point star and half 125-point box, T = 50): every grid of
  the batched graph bitwise equal to the per-grid one-step sweeps (the
  default batch configuration -- star and box grids as concurrent chains --
  and three others covering shared launches and the other tile forms), the
  smallest grids bitwise equal to the oracle, and every cell of every grid
  exact on the linear field.
* FP32 3d2r_star at 512^3, plain and compensated, bitwise equal to the
  oracle's FP32 modes at T = 10; its error against the FP64 sweep at steps
  10/100/1000 within the bound a normalised positive stencil allows.
"""
import sys

import numpy as np
import pytest

from tests.oracle_py import FUSED, KAHAN
from uxxxxxxxxh_amd.workloads import dyadic_coeffs, linear_field, uneven_batch

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def normalised(u, case, seed):
    st = u.builtin(case)
    return st.with_coeffs(u.make_coeffs(st, seed))


TWO_STEP_F64 = ["3d1r_star", "3d2r_star", "3d1r_box", "2d1r_star", "2d2r_star", "2d3r_star", "2d1r_box",
                "2d2r_box"]


def full_ext(case):
    return [8192, 8192] if case.startswith("2d") else [512, 512, 512]


@pytest.mark.parametrize("case", TWO_STEP_F64)
def test_headline_default_bitwise_vs_oracle_512(gpu, orc, case):
    u = gpu
    st = normalised(u, case, 1)
    ext, h = full_ext(case), st.radius
    cfg = u.default_config(ext, h, st)
    log(case, ext, "T=4 default", cfg)
    assert cfg.flags & (u.CFG_TEMPORAL2 | u.CFG_TEMPORALK)
    A0, B0 = orc.fill(ext, h, 1), orc.fill(ext, h, 2)
    dA, dB = u.DeviceArray.from_host(A0), u.DeviceArray.from_host(B0)
    try:
        u.sweep_f64(dA, dB, ext, h, st, 4)
        u.synchronize()
        got = dA.download()
    finally:
        dA.close()
        dB.close()
    orc.iterate(A0, B0, ext, h, st.offsets, st.coeffs, 4, FUSED)
    assert np.array_equal(got, A0)


@pytest.mark.parametrize("case", TWO_STEP_F64)
def test_headline_T100_two_step_equals_one_step(gpu, case):
    """The benchmark's own call (T = 100, default) against one launch per
    step (the kernel the oracle parity tests pin at every size)."""
    u = gpu
    st = normalised(u, case, 1)
    ext, h = full_ext(case), st.radius
    shape = u.grid_shape(ext, h)
    res = []
    for cfg in (None, u.one_step_config(ext, h, st)):
        log(case, ext, "T=100", cfg or u.default_config(ext, h, st))
        A, B = u.DeviceArray(shape, np.float64), u.DeviceArray(shape, np.float64)
        try:
            u.fill_uniform(A, ext, h, 1)
            u.fill_uniform(B, ext, h, 2)
            u.sweep_f64(A, B, ext, h, st, 100, cfg)
            u.synchronize()
            res.append(A.download())
        finally:
            A.close()
            B.close()
    assert np.array_equal(res[0], res[1])


@pytest.mark.parametrize("case,compensated", [("3d1r_star", False), ("3d1r_box", False)] +
                         [(c, k) for c in TWO_STEP_F64 if c.startswith("2d") for k in (False, True)])
def test_f32_two_step_default_vs_oracle_512(gpu, orc, case, compensated):
    """The FP32 defaults that advance two or three steps per launch: the
    plain 3d1r_star and 3d1r_box, and the 2D stencils but the radius-3 box,
    plain and compensated (three steps for the stars)."""
    u = gpu
    st = normalised(u, case, 3)
    ext, h = full_ext(case), st.radius
    dt = u.DT_F32K if compensated else u.DT_F32
    cfg = u.default_config(ext, h, st, dt)
    log(case, "f32k" if compensated else "f32", ext, "T=4", cfg)
    assert cfg.flags & (u.CFG_TEMPORAL2 | u.CFG_TEMPORALK)
    A0, B0 = orc.fill(ext, h, 5, np.float32), orc.fill(ext, h, 6, np.float32)
    dA, dB = u.DeviceArray.from_host(A0), u.DeviceArray.from_host(B0)
    try:
        u.sweep_f32(dA, dB, ext, h, st, 4, compensated)
        u.synchronize()
        got = dA.download()
    finally:
        dA.close()
        dB.close()
    orc.iterate(A0, B0, ext, h, st.offsets, st.coeffs, 4, KAHAN if compensated else FUSED)
    assert np.array_equal(got, A0)


CASES = ["2d1r_star", "2d2r_star", "2d3r_star", "2d1r_box", "2d2r_box", "2d3r_box",
         "3d1r_star", "3d2r_star", "3d3r_star", "3d1r_box", "3d2r_box", "3d3r_box"]


@pytest.mark.parametrize("case", CASES)
def test_linear_field_fixed_point_full_size(gpu, case):
    u = gpu
    st = u.builtin(case)
    st = st.with_coeffs(dyadic_coeffs(st.offsets))
    ext = [8192, 8192] if st.dims == 2 else [512, 512, 512]
    h = st.radius
    f = linear_field(u.grid_shape(ext, h))
    log(case, "linear field T=4", u.default_config(ext, h, st))
    A, B = u.DeviceArray.from_host(f), u.DeviceArray.from_host(f)
    try:
        u.sweep_f64(A, B, ext, h, st, 4)
        u.synchronize()
        assert np.array_equal(A.download(), f), "result grid"
        assert np.array_equal(B.download(), f), "other grid"
    finally:
        A.close()
        B.close()


# ---- the benchmark batch -------------------------------------------------

def _batch_grids(u, field=False):
    """Device grids of the benchmark batch: (A, B, ext, h, stencil) and
    pristine copies (A0, B0) to restore them from."""
    grids, pristine = [], []
    for s in uneven_batch():
        st = u.builtin(s["case"])
        st = st.with_coeffs(dyadic_coeffs(st.offsets) if field else u.make_coeffs(st, s["coeff_seed"]))
        ext, h = s["ext"], st.radius
        shape = u.grid_shape(ext, h)
        if field:
            f = linear_field(shape)
            A, B = u.DeviceArray.from_host(f), u.DeviceArray.from_host(f)
        else:
            A, B = u.DeviceArray(shape, np.float64), u.DeviceArray(shape, np.float64)
            u.fill_uniform(A, ext, h, s["seed_a"])
            u.fill_uniform(B, ext, h, s["seed_b"])
        A0, B0 = u.DeviceArray(shape, np.float64), u.DeviceArray(shape, np.float64)
        u.check(u.lib().uhs_copy_f64(A0.ptr, A.ptr, int(np.prod(shape)), None))
        u.check(u.lib().uhs_copy_f64(B0.ptr, B.ptr, int(np.prod(shape)), None))
        grids.append((A, B, ext, h, st))
        pristine.append((A0, B0))
    u.synchronize()
    return grids, pristine


def _restore(u, grids, pristine):
    for (A, B, ext, h, _), (A0, B0) in zip(grids, pristine):
        n = int(np.prod(A.shape))
        u.check(u.lib().uhs_copy_f64(A.ptr, A0.ptr, n, None))
        u.check(u.lib().uhs_copy_f64(B.ptr, B0.ptr, n, None))
    u.synchronize()


def _close(grids, pristine):
    for (A, B, _, _, _), (A0, B0) in zip(grids, pristine):
        for d in (A, B, A0, B0):
            d.close()


def _batch_configs(u):
    d = u.batch_default_config()
    one = u.batch_default_config()
    one.star_depth = 1
    one.star_tile_x, one.star_tile_y, one.star_tile_z = 64, 8, 64
    one.box_tile_x, one.box_tile_y, one.box_tile_z, one.flags = 32, 8, 64, 0
    four = u.batch_default_config()
    four.star_depth = 4
    four.star_tile_x, four.star_tile_y, four.star_tile_z = 32, 8, 64
    four.box_tile_x, four.box_tile_y, four.box_tile_z, four.flags = 64, 8, 128, 1
    shared = u.batch_default_config()
    shared.star_tile_x, shared.star_tile_y, shared.star_tile_z = 32, 12, 64
    shared.flags = u.BATCH_BOX_ROWS2 | u.BATCH_BOX_PLANES2
    return [("default", d), ("depth1", one), ("depth4", four), ("shared", shared)]


@pytest.mark.timeout(600)
def test_benchmark_batch_bitwise_vs_per_grid_and_oracle(gpu, orc):
    u = gpu
    T = 50
    grids, pristine = _batch_grids(u)
    try:
        b0 = u.Batch(grids, T)
        log("benchmark batch", b0.info())
        b0.run_each()
        u.synchronize()
        ref = [A.download() for A, _, _, _, _ in grids]
        b0.close()
        # The smallest grids of each kind against the oracle (T = 50).
        vol = [int(np.prod(g[2])) for g in grids]
        small = sorted(range(32), key=lambda i: vol[i])[:3] + sorted(range(32, 64), key=lambda i: vol[i])[:2]
        for i in small:
            A0, B0 = pristine[i][0].download(), pristine[i][1].download()
            _, _, ext, h, st = grids[i]
            log("oracle T=50 grid", i, st.name, ext)
            orc.iterate(A0, B0, ext, h, st.offsets, st.coeffs, T, FUSED)
            assert np.array_equal(ref[i], A0), ("per-grid sweeps vs oracle", i)
        for name, cfg in _batch_configs(u):
            _restore(u, grids, pristine)
            b = u.Batch(grids, T, cfg)
            log("batch", name, cfg, b.info())
            b.run()
            u.synchronize()
            b.close()
            for i, (A, _, ext, h, st) in enumerate(grids):
                assert np.array_equal(A.download(), ref[i]), (name, i, st.name, ext)
    finally:
        _close(grids, pristine)


@pytest.mark.timeout(600)
def test_benchmark_batch_linear_field_every_cell(gpu):
    u = gpu
    T = 50
    grids, pristine = _batch_grids(u, field=True)
    try:
        for name, cfg in _batch_configs(u)[:1]:
            b = u.Batch(grids, T, cfg)
            log("batch linear field", name, b.info())
            b.run()
            u.synchronize()
            b.close()
            for i, (A, B, ext, h, st) in enumerate(grids):
                f = linear_field(A.shape)
                assert np.array_equal(A.download(), f), (name, i, st.name, ext, "A")
                assert np.array_equal(B.download(), f), (name, i, st.name, ext, "B")
    finally:
        _close(grids, pristine)


# ---- FP32 3d2r_star at 512^3 ---------------------------------------------

@pytest.mark.parametrize("compensated", [False, True])
def test_f32_3d2r_star_512_bitwise_vs_oracle(gpu, orc, compensated):
    u = gpu
    st = normalised(u, "3d2r_star", 1)
    ext, h = [512, 512, 512], 2
    log("3d2r_star f32 512^3 T=10 compensated=%s" % compensated,
        u.default_config(ext, h, st, u.DT_F32K if compensated else u.DT_F32))
    A0, B0 = orc.fill(ext, h, 1, np.float32), orc.fill(ext, h, 2, np.float32)
    dA, dB = u.DeviceArray.from_host(A0), u.DeviceArray.from_host(B0)
    try:
        u.sweep_f32(dA, dB, ext, h, st, 10, compensated)
        u.synchronize()
        got = dA.download()
    finally:
        dA.close()
        dB.close()
    orc.iterate(A0, B0, ext, h, st.offsets, st.coeffs, 10, KAHAN if compensated else FUSED)
    assert np.array_equal(got, A0)


@pytest.mark.timeout(300)
def test_f32_3d2r_star_512_error_bounds():
    """Errors against the FP64 sweep on identical inputs after 10/100/1000
    steps.  The coefficients are positive and sum to 1, so a sweep does not
    amplify an existing error in the max norm; each step adds at most the
    rounding of its sum (13 terms: 13 u plain, 2 u compensated, |x| < 1) and
    of the stored result (u), u = 2^-24.  Hence max-abs <= T (13+1) u plain and
    T (2+1) u compensated."""
    from uxxxxxxxxh_amd.fp32study import study
    import uxxxxxxxxh_amd as u
    r = study("3d2r_star", [512, 512, 512], [10, 100, 1000])
    uu = 2.0 ** -24
    for T, e in r["checkpoints"].items():
        T = int(T)
        log("3d2r_star f32 step", T, e)
        assert 0 < e["f32"]["max_abs"] <= T * 14 * uu
        assert 0 < e["f32k"]["max_abs"] <= T * 3 * uu
        assert 0 < e["f32"]["rel_l2"] < 1e-3 and 0 < e["f32k"]["rel_l2"] < 1e-3
