"""The allocation-free norm entry (uhs_error_norms_f32_f64_noalloc) against
the allocating one it mirrors, through the C ABI: bitwise the same three
norms, on two calls in a row, the second on a smaller grid (the static
scratch keeps the first call's partial sums, which must not leak into the
second result)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _norms(u, fn, a, b, ext, h):
    out = (ctypes.c_double * 3)()
    u.check(getattr(u.lib(), fn)(a.ptr, b.ptr, len(ext), u.ext_array(ext), h, out, None))
    return tuple(out)


def _grids(u, ext, h, seed):
    shape = u.grid_shape(ext, h)
    b = u.DeviceArray(shape, np.float64)
    u.fill_uniform(b, ext, h, seed)
    host = b.download()
    # FP32 grid: the FP64 values rounded, plus a seeded perturbation.
    rng = np.random.default_rng(seed)
    a = u.DeviceArray.from_host((host + 1e-3 * rng.standard_normal(shape)).astype(np.float32))
    return a, b


@pytest.mark.parametrize("dims", [3, 2])
def test_noalloc_norms_equal_allocating(gpu, dims):
    u = gpu
    cases = ([[96, 80, 130], [17, 9, 33]] if dims == 3 else [[700, 513], [31, 40]])
    for k, ext in enumerate(cases):          # large first, then smaller
        h = 2
        a, b = _grids(u, ext, h, 11 + k)
        try:
            ref = _norms(u, "uhs_error_norms_f32_f64", a, b, ext, h)
            got = _norms(u, "uhs_error_norms_f32_f64_noalloc", a, b, ext, h)
            assert got == ref, (ext, got, ref)
            assert ref[0] > 0 and ref[1] > 0 and ref[2] > 0
            # A closed form on the host for the maximum: |a - b| over the interior.
            ha, hb = a.download().astype(np.float64), b.download()
            inner = tuple(slice(h, h + e) for e in ext)
            assert got[0] == np.max(np.abs(ha[inner] - hb[inner]))
        finally:
            a.close()
            b.close()


def test_noalloc_norms_reject_bad_arguments(gpu):
    u = gpu
    out = (ctypes.c_double * 3)()
    L = u.lib()
    assert L.uhs_error_norms_f32_f64_noalloc(None, None, 3, u.ext_array([4, 4, 4]), 1, out, None) == -1
    assert L.uhs_error_norms_f32_f64_noalloc(None, None, 4, u.ext_array([4, 4, 4]), 1, out, None) == -1
