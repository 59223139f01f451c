"""Pins the CPU oracle (oracle/uhs_oracle.c) before anything is compared with it.

* Known-answer vectors (tests/golden/kat_*.npz) come from evaluating each
  reference DSL `expr` text with Python's own parser, so a bitwise match of
  ORC_PLAIN pins term order, offsets, coefficients, halo handling and
  ping-pong against the reference's own definition (make_fixtures.py).
* Closed forms pin what the vectors leave open for arbitrary coefficients:
  the impulse response is the reflected stencil, a 2^53 + 1 probe pins the
  left-to-right summation, a constant field is preserved when sum(c) = 1.
"""
import json
import os

import numpy as np
import pytest

from tests.oracle_py import FUSED, KAHAN, PLAIN

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = json.load(open(os.path.join(HERE, "golden", "shapes.json")))
CASES = list(SHAPES)


def kat(case):
    return np.load(os.path.join(HERE, "golden", "kat_%s.npz" % case))


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("coefs", ["dsl", "sym"])
def test_kat_plain_bitwise(orc, case, coefs):
    s, v = SHAPES[case], kat(case)
    c = s["coeffs"] if coefs == "dsl" else v["sym_coeffs"]
    ext, r = [int(e) for e in v["interior"]], int(v["halo"])
    for T in (1, 2, 3):
        A, B = v["A0"].copy(), v["B0"].copy()
        orc.iterate(A, B, ext, r, s["offsets"], c, T, PLAIN)
        assert np.array_equal(A, v["%s_A_T%d" % (coefs, T)]), (case, T)
        assert np.array_equal(B, v["%s_B_T%d" % (coefs, T)]), (case, T)


@pytest.mark.parametrize("case", CASES)
def test_kat_fused_within_1e12(orc, case):
    s, v = SHAPES[case], kat(case)
    ext, r = [int(e) for e in v["interior"]], int(v["halo"])
    A, B = v["A0"].copy(), v["B0"].copy()
    orc.iterate(A, B, ext, r, s["offsets"], v["sym_coeffs"], 2, FUSED)
    ref = v["sym_A_T2"]
    assert np.max(np.abs(A - ref) / np.maximum(np.abs(ref), 1e-300)) <= 1e-12


def interior(a, r):
    return a[tuple(slice(r, n - r) for n in a.shape)]


@pytest.mark.parametrize("case", CASES)
def test_impulse_response_is_reflected_stencil(orc, case):
    s = SHAPES[case]
    d, r = s["dims"], s["radius"]
    ext = [9, 11] if d == 2 else [7, 8, 9]
    a = np.zeros([e + 2 * r for e in ext])
    p = tuple(e // 2 + r for e in ext)
    a[p] = 1.0
    b = np.zeros_like(a)
    c = np.arange(1, s["npoints"] + 1) / 64.0  # distinct, exactly representable
    orc.sweep(a, b, ext, r, s["offsets"], c, PLAIN)
    expect = np.zeros_like(a)
    for o, ct in zip(s["offsets"], c):
        expect[tuple(pi - oi for pi, oi in zip(p, o))] = ct
    assert np.array_equal(b, expect)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("mode", [PLAIN, FUSED])
def test_left_to_right_term_order(orc, case, mode):
    """2^53 on the first term absorbs every later +1; on the last term it
    receives the exact sum P-1 first."""
    s = SHAPES[case]
    d, r, P = s["dims"], s["radius"], s["npoints"]
    ext = [2 * r + 3] * d
    p = tuple(e // 2 + r for e in ext)
    for big_term, expect in ((0, 2.0 ** 53), (P - 1, float(2 ** 53 + P - 1))):
        a = np.zeros([e + 2 * r for e in ext])
        for t, o in enumerate(s["offsets"]):
            a[tuple(pi + oi for pi, oi in zip(p, o))] = 2.0 ** 53 if t == big_term else 1.0
        b = np.zeros_like(a)
        orc.sweep(a, b, ext, r, s["offsets"], np.ones(P), mode)
        assert b[p] == expect, (case, big_term, b[p], expect)


def _residual(c, v, dt):
    """c*v - round(c*v), exactly (the error of a rounded product is
    representable in the same format, so the conversion is exact)."""
    from fractions import Fraction
    prod = dt(c) * dt(v)
    r = Fraction(float(dt(c))) * Fraction(float(dt(v))) - Fraction(float(prod))
    assert r != 0 and float(dt(float(r))) == r
    return float(r), float(prod)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_fused_mode_contracts(orc, case, dt):
    """FUSED evaluates term t as one fused multiply-add; PLAIN rounds the
    product first.  Term 0 contributes s = -round(c1*v) exactly (coefficient
    1), term 1 then adds c1*v with c1*v inexact: fma(c1, v, s) is the exact
    rounding error of the product (nonzero), the unfused sum is 0.  Later
    terms read zeros.  A silently unfused FUSED mode, or a contracted PLAIN
    mode, fails here."""
    s = SHAPES[case]
    d, r, P = s["dims"], s["radius"], s["npoints"]
    ext = [2 * r + 3] * d
    p = tuple(e // 2 + r for e in ext)
    c1, v = 0.1, 3.0
    res, prod = _residual(c1, v, dt)
    a = np.zeros([e + 2 * r for e in ext], dt)
    off = s["offsets"]
    a[tuple(pi + oi for pi, oi in zip(p, off[0]))] = -prod
    a[tuple(pi + oi for pi, oi in zip(p, off[1]))] = v
    c = np.full(P, 0.5)
    c[0], c[1] = 1.0, c1
    for mode, expect in ((FUSED, res), (PLAIN, 0.0)):
        b = np.zeros_like(a)
        orc.sweep(a, b, ext, r, off, c, mode)
        assert float(b[p]) == expect, (case, mode, float(b[p]), expect)


@pytest.mark.parametrize("case", CASES)
def test_constant_field_preserved(orc, case):
    s = SHAPES[case]
    d, r = s["dims"], s["radius"]
    ext = [12, 10] if d == 2 else [6, 5, 7]
    c = orc.make_coeffs(s["offsets"], d, 5)
    a = np.full([e + 2 * r for e in ext], 0.625)
    b = np.zeros_like(a)
    orc.sweep(a, b, ext, r, s["offsets"], c, PLAIN)
    assert np.max(np.abs(interior(b, r) - 0.625)) <= 4 * s["npoints"] * 2.0 ** -53
    assert np.all(b[tuple(slice(0, r) for _ in range(d))] == 0)  # halo untouched


@pytest.mark.parametrize("case", ["2d3r_box", "3d2r_star", "3d3r_box"])
def test_linearity(orc, case):
    s = SHAPES[case]
    d, r = s["dims"], s["radius"]
    ext = [14, 13] if d == 2 else [6, 7, 5]
    c = orc.make_coeffs(s["offsets"], d, 9)
    x, y = orc.fill(ext, r, 1), orc.fill(ext, r, 2)
    outs = []
    for g in (x, y, 3.0 * x + y):
        o = np.zeros_like(g)
        orc.sweep(np.ascontiguousarray(g), o, ext, r, s["offsets"], c, FUSED)
        outs.append(o)
    assert np.allclose(outs[2], 3.0 * outs[0] + outs[1], rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("case", ["2d1r_star", "3d1r_box"])
def test_pingpong_is_repeated_sweep_and_threads_invariant(orc, case):
    s = SHAPES[case]
    d, r = s["dims"], s["radius"]
    ext = [33, 17] if d == 2 else [9, 10, 11]
    c = orc.make_coeffs(s["offsets"], d, 3)
    A0 = orc.fill(ext, r, 11)
    B0 = orc.fill(ext, r, 12)
    A, B = A0.copy(), B0.copy()
    orc.iterate(A, B, ext, r, s["offsets"], c, 4, FUSED, nthreads=1)
    a, b = A0.copy(), B0.copy()
    for _ in range(2):
        orc.sweep(a, b, ext, r, s["offsets"], c, FUSED, nthreads=3)
        orc.sweep(b, a, ext, r, s["offsets"], c, FUSED, nthreads=5)
    assert np.array_equal(A, a) and np.array_equal(B, b)


def test_splitmix64_known_answer(orc):
    # splitmix64 seeded with 0 yields 0xE220A8397B1DCDAF first (Steele et al.,
    # the published reference output); fill and coefficients use this mixer.
    assert orc.mix64(0x9E3779B97F4A7C15) == 0xE220A8397B1DCDAF


def test_fill_uniform(orc):
    ext, h = [20, 21, 22], 2
    a = orc.fill(ext, h, 42)
    inn = interior(a, h)
    assert np.all((inn >= 0) & (inn < 1)) and abs(inn.mean() - 0.5) < 0.02
    mask = np.ones(a.shape, bool)
    mask[tuple(slice(h, n - h) for n in a.shape)] = False
    assert np.all(a[mask] == 0)
    L = (h * a.shape[1] + h) * a.shape[2] + h  # first interior cell
    z = orc.mix64((42 + (L + 1) * 0x9E3779B97F4A7C15) & (2 ** 64 - 1))
    assert a.reshape(-1)[L] == (z >> 11) * 2.0 ** -53
    assert np.array_equal(orc.fill(ext, h, 42, np.float32), a.astype(np.float32))


@pytest.mark.parametrize("case", CASES)
def test_coeffs_symmetric_normalised(orc, case):
    s = SHAPES[case]
    c = orc.make_coeffs(s["offsets"], s["dims"], 1234)
    assert abs(c.sum() - 1.0) < 1e-14 and np.all(c > 0)
    idx = {tuple(o): t for t, o in enumerate(s["offsets"])}
    for t, o in enumerate(s["offsets"]):
        assert c[idx[tuple(-v for v in o)]] == c[t]


def test_kahan_beats_plain_f32(orc):
    s = SHAPES["3d3r_box"]
    ext, r = [8, 8, 8], 3
    c = orc.make_coeffs(s["offsets"], 3, 7)
    x = orc.fill(ext, r, 5) + 1000.0 * (orc.fill(ext, r, 6) > 0.5)
    x = np.ascontiguousarray(x)
    ref = np.zeros_like(x)
    orc.sweep(x, ref, ext, r, s["offsets"], c, FUSED)
    errs = {}
    for mode in (FUSED, KAHAN):
        x32 = x.astype(np.float32)
        o = np.zeros_like(x32)
        orc.sweep(x32, o, ext, r, s["offsets"], c, mode)
        errs[mode] = np.abs(interior(o, r) - interior(ref, r)).max()
    assert errs[KAHAN] < errs[FUSED]


@pytest.mark.parametrize("case", ["2d3r_box", "3d2r_box", "3d3r_box"])
def test_kahan_recovers_absorbed_terms(orc, case):
    """1.0 followed by P-1 terms of 2^-25: in FP32 every plain addition of
    2^-25 to 1.0 is absorbed (below half an ulp), so the plain sum is exactly
    1; the compensated sum recovers 1 + (P-1) * 2^-25 to within one rounding."""
    s = SHAPES[case]
    d, r, P = s["dims"], s["radius"], s["npoints"]
    ext = [2 * r + 3] * d
    p = tuple(e // 2 + r for e in ext)
    a = np.zeros([e + 2 * r for e in ext], np.float32)
    for t, o in enumerate(s["offsets"]):
        a[tuple(pi + oi for pi, oi in zip(p, o))] = 1.0 if t == 0 else 2.0 ** -25
    exact = 1.0 + (P - 1) * 2.0 ** -25
    for mode, expect_exact in ((FUSED, False), (KAHAN, True)):
        b = np.zeros_like(a)
        orc.sweep(a, b, ext, r, s["offsets"], np.ones(P), mode)
        if expect_exact:
            assert abs(float(b[p]) - exact) <= 2.0 ** -23
        else:
            assert b[p] == np.float32(1.0)
