"""Host-side checks of libuhs.so that need no GPU: exported symbols, the
built-in stencil table, coefficient generation and argument validation."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SHAPES = json.load(open(os.path.join(HERE, "golden", "shapes.json")))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(uhs_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(uhs):
    L = uhs.lib()
    names = declared("uhs.h")
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_builtin_table_matches_dsl_tables(uhs):
    names = uhs.builtin_names()
    assert names == list(SHAPES)
    for case in names:
        st, s = uhs.builtin(case), SHAPES[case]
        assert st.symbol == s["symbol"] and st.dims == s["dims"] and st.radius == s["radius"]
        assert st.offsets.tolist() == s["offsets"]
        assert st.coeffs.tolist() == s["coeffs"]


@pytest.mark.parametrize("case", list(SHAPES))
def test_make_coeffs_matches_oracle(uhs, orc, case):
    st = uhs.builtin(case)
    assert np.array_equal(uhs.make_coeffs(st, 77), orc.make_coeffs(st.offsets, st.dims, 77))


def test_sym_fixture_coefficients_match(uhs):
    st = uhs.builtin("3d2r_box")
    v = np.load(os.path.join(HERE, "golden", "kat_3d2r_box.npz"))
    assert np.array_equal(uhs.make_coeffs(st, 77), v["sym_coeffs"])


@pytest.mark.parametrize("case", list(SHAPES))
def test_plan_kinds_and_default_configs(uhs, case):
    s = SHAPES[case]
    st = uhs.builtin(case)
    ext = [8192, 8192] if s["dims"] == 2 else [512, 512, 512]
    kind = uhs.plan_kind(ext, s["radius"], st)
    assert kind == ("win2d" if s["dims"] == 2 else ("star3d" if s["kind"] == "star" else "box3d"))
    for dt in (uhs.DT_F64, uhs.DT_F32, uhs.DT_F32K):
        c = uhs.default_config(ext, s["radius"], st, dt)
        # The measured default is compiled for every precision: the
        # fallback for a configuration without a kernel (vec 1, no flags)
        # is never taken at the benchmark sizes (no table entry is vec 1
        # without flags).
        assert uhs.plan_kind(ext, s["radius"], st, dt, cfg=c.tuple()) == kind
        assert c.vec > 1 or c.flags != 0, (case, dt, c.tuple())
    small = [5, 7] if s["dims"] == 2 else [3, 5, 7]
    uhs.default_config(small, s["radius"], st)


# The 2D multi-step defaults at 8192^2 (tile_x, tile_y, tile_z, vec, flags):
# the stars UHS_CFG_TEMPORALK (256, K = tile_y), the boxes UHS_CFG_TEMPORAL2 (32), with
# UHS_CFG_NT_STORE (1) and, for the window form, UHS_CFG_ROWRING (128).
_T2, _NT, _RR, _T3 = 32, 1, 128, 256
TWO_STEP_2D = {
    0: {"2d1r_star": (242, 8, 64, 1, _T3 | _NT), "2d2r_star": (236, 6, 64, 1, _T3),
        "2d3r_star": (238, 4, 64, 1, _T3), "2d1r_box": (246, 6, 64, 1, _T3 | _NT),
        "2d2r_box": (244, 4, 64, 1, _T3 | _NT)},
    1: {"2d1r_star": (242, 8, 64, 1, _T3), "2d2r_star": (236, 6, 64, 1, _T3),
        "2d3r_star": (488, 5, 64, 1, _T3), "2d1r_box": (118, 6, 64, 1, _T3 | _NT),
        "2d2r_box": (252, 1, 64, 1, _T2 | _RR)},
    2: {"2d1r_star": (242, 8, 64, 1, _T3 | _NT), "2d2r_star": (236, 6, 64, 1, _T3),
        "2d3r_star": (488, 5, 64, 1, _T3 | _NT), "2d1r_box": (250, 4, 64, 1, _T3 | _NT),
        "2d2r_box": (124, 1, 64, 1, _T2 | _NT | _RR)},
}


@pytest.mark.parametrize("case", list(SHAPES))
def test_one_step_configs(uhs, case):
    """Defaults that advance two steps per launch: the FP64 radius-1 and
    radius-2 3D stars, the plain FP32 radius-1 3D star, the FP64 and plain
    FP32 radius-1 3D box, and the plain and compensated FP32 radius-2 2D
    box; K steps per launch (K = tile_y): the 2D stars and the radius-1 2D
    box in every precision and the FP64 radius-2 box.  Their one-step configuration
    (leading steps, unbatched baseline) is the one-step default tile; every
    other default is its own one-step configuration."""
    s = SHAPES[case]
    st = uhs.builtin(case)
    for ext in ([8192, 8192] if s["dims"] == 2 else [512, 512, 512],
                [64, 96] if s["dims"] == 2 else [64, 96, 128]):
        for dt in (uhs.DT_F64, uhs.DT_F32, uhs.DT_F32K):
            d = uhs.default_config(ext, s["radius"], st, dt)
            o = uhs.one_step_config(ext, s["radius"], st, dt)
            assert not o.flags & (uhs.CFG_TEMPORAL2 | uhs.CFG_TEMPORALK)
            assert uhs.plan_kind(ext, s["radius"], st, dt, cfg=o.tuple()) == uhs.plan_kind(ext, s["radius"], st, dt)
            big = ext[0] == 512
            if case == "3d1r_star" and dt in (uhs.DT_F64, uhs.DT_F32):
                assert d.tuple() == (62, 14, 32 if big else 8, 2, uhs.CFG_TEMPORAL2)
                if dt == uhs.DT_F64:
                    assert o.tuple() == (128, 16, 256 if big else 8, 8, uhs.CFG_NT_STORE)
                else:
                    assert o.tuple()[:2] == (256, 16) and o.flags == 0
            elif case == "3d1r_box" and dt in (uhs.DT_F64, uhs.DT_F32):
                assert d.tuple()[:2] == ((62, 14) if dt == uhs.DT_F64 else (30, 14)) and d.vec == 2
                assert d.flags == uhs.CFG_TEMPORAL2 | uhs.CFG_NT_STORE
                assert not o.flags & uhs.CFG_TEMPORAL2 and o.flags & uhs.CFG_ROWS2
            elif case == "3d2r_star" and dt == uhs.DT_F64:
                assert d.tuple() == (28, 12, 64 if big else 16, 2, uhs.CFG_TEMPORAL2 | uhs.CFG_NT_STORE)
                assert o.tuple() == (128, 16, 256 if big else 16, 8, uhs.CFG_NT_STORE)
            elif s["dims"] == 2 and case != "2d3r_box":
                tb = TWO_STEP_2D[dt][case]
                assert d.flags == tb[4] and (d.tile_x, d.tile_y, d.vec) == (tb[0], tb[1], tb[3])
                assert d.tile_z == tb[2] if ext[0] == 8192 else d.tile_z <= tb[2]
                assert not o.flags & uhs.CFG_TEMPORAL2
            else:
                assert d.tuple() == o.tuple()


def expect_code(uhs, code, fn):
    with pytest.raises(uhs.UhsError) as e:
        fn()
    assert e.value.code == code, str(e.value)


def test_argument_validation(uhs):
    st = uhs.builtin("3d1r_star")
    ext = [64, 64, 64]
    expect_code(uhs, -1, lambda: uhs.plan_kind(ext, 0, st))             # halo < radius
    expect_code(uhs, -1, lambda: uhs.plan_kind([64, 64], 1, st))        # dims mismatch
    expect_code(uhs, -1, lambda: uhs.plan_kind([0, 64, 64], 1, st))     # empty extent
    expect_code(uhs, -3, lambda: uhs.plan_kind(ext, 1, st, cfg=(64, 8, 64, 3)))
    expect_code(uhs, -3, lambda: uhs.plan_kind(ext, 1, st, cfg=(1024, 8, 64, 1)))
    expect_code(uhs, -3, lambda: uhs.plan_kind([64, 64], 1, uhs.builtin("2d1r_star"), cfg=(64, 2, 8, 1)))
    # the row ring: 2D only, not with the shifted prefetch
    s2 = uhs.builtin("2d1r_star")
    assert uhs.plan_kind([64, 64], 1, s2, cfg=(64, 1, 8, 1, uhs.CFG_ROWRING)) == "win2d"
    expect_code(uhs, -3, lambda: uhs.plan_kind([64, 64], 1, s2, cfg=(64, 1, 8, 1, uhs.CFG_ROWRING | uhs.CFG_PREFETCH2)))
    expect_code(uhs, -3, lambda: uhs.plan_kind(ext, 1, st, cfg=(64, 8, 64, 1, uhs.CFG_ROWRING)))
    big = uhs.StencilSpec(3, [[i % 5 - 2, (i // 5) % 5 - 2, i // 25 - 2] for i in range(125)][::-1],
                          np.ones(125))
    assert uhs.plan_kind(ext, 2, big) == "gather" or True
    many = [[a, b, c] for a in range(-3, 4) for b in range(-3, 4) for c in range(-3, 4)][::-1]
    expect_code(uhs, -2, lambda: uhs.plan_kind(ext, 3, uhs.StencilSpec(3, many, np.ones(343))))
    expect_code(uhs, -2, lambda: uhs.plan_kind(ext, 5, uhs.StencilSpec(3, [[5, 0, 0]], [1.0])))
    assert uhs.plan_kind(ext, 2, uhs.StencilSpec(3, [[0, 0, 0], [1, 0, 0]], [0.5, 0.5])) == "program"
    assert uhs.plan_kind(ext, 2, uhs.StencilSpec(3, [[0, 0, 0], [1, 0, 0]], [0.5, 0.5]), uhs.DT_F32) == "program"


def test_probe_argument_checks(uhs):
    # Rejected before any device call.
    L = uhs.lib()
    assert L.uhs_fma_f64(None, 10, 1.0, 0.0, 1, None) == -1
    assert L.uhs_copy_f64(None, None, 10, None) == -1


def test_model_search_argument_checks(uhs):
    # The ensemble search rejects a missing model list, an empty ensemble, a
    # null member and topk < 1 before planning or any device call.
    import ctypes
    L = uhs.lib()
    st = uhs.builtin("2d3r_star")
    ext = (ctypes.c_int64 * 2)(6144, 12288)
    r = uhs.TuneResult()
    none = (ctypes.c_void_p * 2)(None, None)
    for models, n, topk in ((None, 1, 5), (none, 0, 5), (none, 2, 5), (none, 1, 0)):
        rc = L.uhs_tune_sweep_models(None, None, 2, ext, 3, ctypes.byref(st.c), uhs.DT_F64, models, n, topk, 1,
                                     None, ctypes.byref(r), None)
        assert rc == -1, (n, topk)
    assert L.uhs_tune_sweep_model(None, None, 2, ext, 3, ctypes.byref(st.c), uhs.DT_F64, None, 5, 1, None,
                                  ctypes.byref(r), None) == -1


def test_no_gpu_needed_for_planning(uhs):
    # Planning, defaults and coefficient generation are host-only.
    st = uhs.builtin("2d3r_box")
    c = uhs.default_config([8192, 8192], 3, st)
    assert c.tile_y == 1 and c.tile_x % c.vec == 0


def test_two_step_config_checks_host_only(uhs):
    """The two-step forms' legality rules, decided on the host without a
    GPU: which stencils and precisions have a two-step kernel, and the
    limits on columns per thread, threads and flags."""
    T2, NT, RR, PF2 = uhs.CFG_TEMPORAL2, uhs.CFG_NT_STORE, uhs.CFG_ROWRING, uhs.CFG_PREFETCH2
    F64, F32 = uhs.DT_F64, uhs.DT_F32
    s1, s3, b1, b3 = (uhs.builtin(c) for c in ("2d1r_star", "2d3r_star", "2d1r_box", "2d3r_box"))
    e2, e3 = [64, 96], [40, 44, 70]
    ok = [(e2, s1, F64, (254, 1, 32, 4, T2)),            # 2D star, ring form
          (e2, s1, F64, (254, 1, 32, 4, T2 | RR)),       # 2D star, window form
          (e2, s3, F64, (58, 1, 32, 4, T2)),
          (e2, s3, F64, (58, 1, 32, 1, T2 | RR | NT)),
          (e2, b1, F64, (62, 1, 32, 2, T2 | RR)),
          (e2, b3, F32, (58, 1, 32, 1, T2 | RR)),        # FP32 radius-3 box, window form
          (e2, b1, F64, (62, 1, 32, 2, T2)),             # 2D box, sum form
          (e2, b3, F64, (58, 1, 32, 1, T2 | NT)),        # FP64 radius-3 box, sum form
          (e3, uhs.builtin("3d1r_box"), F64, (62, 14, 32, 2, T2 | NT)),
          (e3, uhs.builtin("3d1r_box"), F32, (14, 6, 8, 4, T2)),
          (e3, uhs.builtin("3d1r_star"), F64, (62, 14, 32, 2, T2))]
    T3 = uhs.CFG_TEMPORALK
    ok += [(e2, b1, F64, (60, 3, 32, 2, T3)),                 # K steps, radius-1 box (window form)
           (e2, s1, F64, (252, 3, 32, 2, T3)),                # K = tile_y steps per launch, 2D stars
           (e2, s1, F64, (242, 8, 32, 2, T3)),
           (e2, s3, F32, (52, 3, 32, 2, T3 | NT))]
    for ext, st, dt, cfg in ok:
        assert uhs.plan_kind(ext, st.radius, st, dt, cfg=cfg) in ("win2d", "box3d", "star3d"), cfg
    bad = [(e2, s1, F64, (61, 1, 32, 2, T2)),            # tile_x + 2r odd with vec 2
           (e2, s1, F64, (62, 2, 32, 2, T2)),            # tile_y 2 in 2D
           (e2, s1, F64, (1022, 1, 32, 1, T2)),          # 1024 threads
           (e2, s1, F64, (62, 1, 32, 2, T2 | PF2)),      # no deeper prefetch in 2D
           (e2, s3, F64, (58, 1, 32, 2, T2 | RR)),       # FP64 radius-3 window: one column per thread
           (e2, b1, F64, (62, 1, 32, 4, T2)),            # sum form: 1 or 2 columns per thread
           (e2, b3, F64, (58, 1, 32, 1, T2 | RR)),       # no FP64 radius-3 box window form
           (e3, uhs.builtin("3d1r_box"), F64, (14, 6, 8, 4, T2)),          # FP64 box: vec <= 2
           (e3, uhs.builtin("3d1r_box"), F64, (62, 14, 32, 2, T2 | RR)),   # box flags: NT only
           (e3, uhs.builtin("3d2r_box"), F64, (28, 12, 32, 1, T2)),        # radius-2 box: none
           (e3, uhs.builtin("3d1r_box"), F64, (126, 14, 32, 2, T2)),       # 1024 threads
           (e2, s1, F64, (62, 3, 32, 2, T3 | T2)),       # K steps: not with TEMPORAL2
           (e2, s1, F64, (60, 3, 32, 2, T3 | RR)),       # K steps: ring form only
           (e2, s1, F64, (62, 1, 32, 2, T3)),            # K steps: K = tile_y of 3, 4, 5, 6 or 8
           (e2, s1, F64, (58, 7, 32, 2, T3)),
           (e2, s3, F64, (52, 3, 32, 4, T3)),            # K steps: 4 columns for the radius-1 star only
           (e2, b3, F64, (52, 3, 32, 1, T3)),            # K steps: not the radius-3 box
           (e3, uhs.builtin("3d1r_star"), F64, (60, 3, 32, 2, T3))]  # K steps: 2D only
    for ext, st, dt, cfg in bad:
        with pytest.raises(uhs.UhsError) as e:
            uhs.plan_kind(ext, st.radius, st, dt, cfg=cfg)
        assert e.value.code == -3, cfg
