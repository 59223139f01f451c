"""MI355X (gfx950) stencil sweep: Python bindings over the C ABI in libuhs.so.

The product is the shared library built from ``csrc/`` (HIP kernels and the
C ABI declared in ``include/uhs.h``).  This module only loads it with ctypes
and wraps device buffers and stencil descriptions for tests and the
benchmark.  Nothing here computes a sweep on the CPU: if ``libuhs.so`` is
missing, :func:`lib` raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libuhs.so")

UHS_OK = 0
ERRORS = {-1: "UHS_E_INVAL", -2: "UHS_E_UNSUPPORTED", -3: "UHS_E_CONFIG", -4: "UHS_E_HIP",
          -5: "UHS_E_NOMEM", -6: "UHS_E_IO", -7: "UHS_E_PARSE"}
DT_F64, DT_F32, DT_F32K = 0, 1, 2
CFG_NT_STORE, CFG_PREFETCH2, CFG_REGTILE, CFG_PREFETCH4, CFG_ROWS2, CFG_TEMPORAL2, CFG_PLANES2 = 1, 2, 4, 8, 16, 32, 64
CFG_ROWRING = 128
CFG_TEMPORALK = 256
BATCH_BOX_ROWS2, BATCH_BOX_PLANES2, BATCH_SPLIT, BATCH_CONCURRENT = 1, 2, 4, 8


class UhsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class Stencil(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int), ("npoints", ctypes.c_int),
                ("offsets", ctypes.POINTER(ctypes.c_int)),
                ("coeffs", ctypes.POINTER(ctypes.c_double))]


MAX_PRODUCERS, MAX_PROGRAM_TERMS = 4, 32


class Program(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int), ("nproducers", ctypes.c_int), ("producers", ctypes.POINTER(Stencil)),
                ("output", Stencil), ("sources", ctypes.POINTER(ctypes.c_int)),
                ("lo", ctypes.c_int64 * 3), ("hi", ctypes.c_int64 * 3),
                ("producer_sources", ctypes.POINTER(ctypes.c_int))]


class DslProgram(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("operation", ctypes.c_char * 64), ("dims", ctypes.c_int),
                ("is_float", ctypes.c_int), ("iteration", ctypes.c_int), ("grid", ctypes.c_int64 * 3),
                ("domain_lo", ctypes.c_int64 * 3), ("domain_hi", ctypes.c_int64 * 3), ("nproducers", ctypes.c_int),
                ("producer_names", (ctypes.c_char * 64) * MAX_PRODUCERS),
                ("producer_npoints", ctypes.c_int * MAX_PRODUCERS),
                ("producer_offsets", (ctypes.c_int * (MAX_PROGRAM_TERMS * 3)) * MAX_PRODUCERS),
                ("producer_coeffs", (ctypes.c_double * MAX_PROGRAM_TERMS) * MAX_PRODUCERS),
                ("producer_sources", (ctypes.c_int * MAX_PROGRAM_TERMS) * MAX_PRODUCERS),
                ("npoints", ctypes.c_int), ("offsets", ctypes.c_int * (MAX_PROGRAM_TERMS * 3)),
                ("coeffs", ctypes.c_double * MAX_PROGRAM_TERMS), ("sources", ctypes.c_int * MAX_PROGRAM_TERMS)]


class Config(ctypes.Structure):
    _fields_ = [("tile_x", ctypes.c_int), ("tile_y", ctypes.c_int),
                ("tile_z", ctypes.c_int), ("vec", ctypes.c_int), ("flags", ctypes.c_int)]

    def __repr__(self):
        return "Config(%d,%d,%d,%d,%d)" % (self.tile_x, self.tile_y, self.tile_z, self.vec, self.flags)

    def tuple(self):
        return (self.tile_x, self.tile_y, self.tile_z, self.vec, self.flags)


class DslInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("kernel", ctypes.c_char * 64), ("dims", ctypes.c_int),
                ("is_float", ctypes.c_int), ("iteration", ctypes.c_int), ("halo", ctypes.c_int),
                ("npoints", ctypes.c_int), ("grid", ctypes.c_int64 * 3), ("domain_lo", ctypes.c_int64 * 3),
                ("domain_hi", ctypes.c_int64 * 3), ("extents", ctypes.c_int64 * 3),
                ("offsets", ctypes.c_int * (343 * 3)), ("coeffs", ctypes.c_double * 343)]


class GridDesc(ctypes.Structure):
    _fields_ = [("A", ctypes.c_void_p), ("B", ctypes.c_void_p), ("dims", ctypes.c_int),
                ("ext", ctypes.c_int64 * 3), ("halo", ctypes.c_int), ("stencil", Stencil)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("star_depth", ctypes.c_int), ("star_tile_x", ctypes.c_int), ("star_tile_y", ctypes.c_int),
                ("star_tile_z", ctypes.c_int), ("box_tile_x", ctypes.c_int), ("box_tile_y", ctypes.c_int),
                ("box_tile_z", ctypes.c_int), ("flags", ctypes.c_int)]

    def __repr__(self):
        return "BatchConfig(depth=%d star=%dx%dx%d box=%dx%dx%d flags=%d)" % (
            self.star_depth, self.star_tile_x, self.star_tile_y, self.star_tile_z,
            self.box_tile_x, self.box_tile_y, self.box_tile_z, self.flags)


class TuneResult(ctypes.Structure):
    _fields_ = [("best", Config), ("best_ms", ctypes.c_double), ("evaluated", ctypes.c_int),
                ("failed", ctypes.c_int), ("reused", ctypes.c_int), ("seconds", ctypes.c_double)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_I64P = ctypes.POINTER(ctypes.c_int64)
_SIGS = {
    "uhs_sweep_f64": (_I, [_P, _P, _I, _I64P, _I, ctypes.POINTER(Stencil), _I, ctypes.POINTER(Config), _P]),
    "uhs_sweep_f32": (_I, [_P, _P, _I, _I64P, _I, ctypes.POINTER(Stencil), _I, _I, ctypes.POINTER(Config), _P]),
    "uhs_plan_kind": (_I, [_I, _I64P, _I, ctypes.POINTER(Stencil), _I, ctypes.POINTER(Config), ctypes.c_char_p, ctypes.c_size_t]),
    "uhs_default_config": (_I, [_I, _I64P, _I, ctypes.POINTER(Stencil), _I, ctypes.POINTER(Config)]),
    "uhs_one_step_config": (_I, [_I, _I64P, _I, ctypes.POINTER(Stencil), _I, ctypes.POINTER(Config)]),
    "uhs_builtin_count": (_I, []),
    "uhs_builtin_info": (_I, [_I, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(_I),
                              ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(ctypes.POINTER(_I)),
                              ctypes.POINTER(ctypes.POINTER(ctypes.c_double))]),
    "uhs_builtin_find": (_I, [ctypes.c_char_p]),
    "uhs_fill_uniform_f64": (_I, [_P, _I, _I64P, _I, ctypes.c_uint64, _P]),
    "uhs_fill_uniform_f32": (_I, [_P, _I, _I64P, _I, ctypes.c_uint64, _P]),
    "uhs_make_coeffs": (_I, [_I, _I, ctypes.POINTER(_I), ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]),
    "uhs_error_norms_f32_f64": (_I, [_P, _P, _I, _I64P, _I, ctypes.POINTER(ctypes.c_double), _P]),
    "uhs_error_norms_f32_f64_noalloc": (_I, [_P, _P, _I, _I64P, _I, ctypes.POINTER(ctypes.c_double), _P]),
    "uhs_error_norms_f64": (_I, [_P, _P, _I, _I64P, _I, ctypes.POINTER(ctypes.c_double), _P]),
    "uhs_parse_dsl": (_I, [ctypes.c_char_p, ctypes.POINTER(DslInfo)]),
    "uhs_parse_dsl_program": (_I, [ctypes.c_char_p, ctypes.POINTER(DslProgram)]),
    "uhs_parse_dsl_program_file": (_I, [ctypes.c_char_p, ctypes.POINTER(DslProgram)]),
    "uhs_program_sweep_f64": (_I, [_P, _P, _I64P, ctypes.POINTER(Program), _I, _P]),
    "uhs_dsl_program_view": (_I, [ctypes.POINTER(DslProgram), ctypes.POINTER(Stencil), ctypes.POINTER(Program)]),
    "uhs_tune_program_f64": (_I, [_P, _P, _I64P, ctypes.POINTER(Program), _I, ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), _P]),
    "uhs_program_sweep_ex_f64": (_I, [_P, _P, _I64P, ctypes.POINTER(Program), _I, _I, ctypes.POINTER(ctypes.c_int), _P]),
    "uhs_program_sweep_f32": (_I, [_P, _P, _I64P, ctypes.POINTER(Program), _I, _I, _P]),
    "uhs_program_sweep_ex_f32": (_I, [_P, _P, _I64P, ctypes.POINTER(Program), _I, _I, _I,
                                      ctypes.POINTER(ctypes.c_int), _P]),
    "uhs_parse_dsl_file": (_I, [ctypes.c_char_p, ctypes.POINTER(DslInfo)]),
    "uhs_batch_create": (_I, [ctypes.POINTER(GridDesc), _I, _I, ctypes.POINTER(BatchConfig), ctypes.POINTER(_P)]),
    "uhs_batch_run": (_I, [_P, _P]),
    "uhs_batch_destroy": (_I, [_P]),
    "uhs_batch_info": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_I),
                            ctypes.POINTER(ctypes.c_size_t)]),
    "uhs_batch_default_config": (_I, [ctypes.POINTER(BatchConfig)]),
    "uhs_sweep_each_f64": (_I, [ctypes.POINTER(GridDesc), _I, _I, _P]),
    "uhs_batch_partition": (_I, [ctypes.POINTER(GridDesc), _I, _I, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_double)]),
    "uhs_tune_sweep": (_I, [_P, _P, _I, _I64P, _I, ctypes.POINTER(Stencil), _I, _I, ctypes.c_char_p,
                            ctypes.POINTER(TuneResult), _P]),
    "uhs_tune_sweep_model": (_I, [_P, _P, _I, _I64P, _I, ctypes.POINTER(Stencil), _I, _P, _I, _I, ctypes.c_char_p,
                                  ctypes.POINTER(TuneResult), _P]),
    "uhs_tune_sweep_models": (_I, [_P, _P, _I, _I64P, _I, ctypes.POINTER(Stencil), _I, ctypes.POINTER(_P), _I, _I, _I,
                                   ctypes.c_char_p, ctypes.POINTER(TuneResult), _P]),
    "uhs_tune_sweep_list": (_I, [_P, _P, _I, _I64P, _I, ctypes.POINTER(Stencil), _I, ctypes.POINTER(Config), _I, _I,
                                 ctypes.c_char_p, ctypes.POINTER(TuneResult), _P]),
    "uhs_batch_space": (_I, [ctypes.POINTER(BatchConfig), _I]),
    "uhs_batch_tune_list": (_I, [ctypes.POINTER(GridDesc), _I, _I, ctypes.POINTER(BatchConfig), _I,
                                 ctypes.POINTER(BatchConfig), ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_double), _P]),
    "uhs_batch_features": (_I, [ctypes.POINTER(GridDesc), _I, ctypes.POINTER(BatchConfig),
                                ctypes.POINTER(ctypes.c_float)]),
    "uhs_batch_tune_model": (_I, [ctypes.POINTER(GridDesc), _I, _I, _P, _I, ctypes.POINTER(BatchConfig),
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), _P]),
    "uhs_batch_tune": (_I, [ctypes.POINTER(GridDesc), _I, _I, ctypes.POINTER(BatchConfig),
                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(BatchConfig), ctypes.POINTER(ctypes.c_double),
                            _I, ctypes.POINTER(_I), _P]),
    "uhs_tune_space": (_I, [_I, _I64P, _I, ctypes.POINTER(Stencil), _I, ctypes.POINTER(Config), _I]),
    "uhs_tune_time": (_I, [_P, _P, _I, _I64P, _I, ctypes.POINTER(Stencil), _I, ctypes.POINTER(Config), _I,
                           ctypes.POINTER(ctypes.c_double), _P]),
    "uhs_tune_load_db": (_I, [ctypes.c_char_p]),
    "uhs_tune_clear": (_I, []),
    "uhs_tune_set_timer": (_I, [_P, _P]),
    "uhs_set_lds_poison": (_I, [_I]),
    "uhs_mlp_features": (_I, [_I, _I, _I, _I, _I64P, ctypes.POINTER(Config), _I, ctypes.POINTER(ctypes.c_float)]),
    "uhs_mlp_create": (_I, [ctypes.c_uint64, ctypes.POINTER(_P)]),
    "uhs_mlp_destroy": (_I, [_P]),
    "uhs_mlp_num_params": (_I, []),
    "uhs_mlp_get_params": (_I, [_P, ctypes.POINTER(ctypes.c_float), _I]),
    "uhs_mlp_set_params": (_I, [_P, ctypes.POINTER(ctypes.c_float), _I]),
    "uhs_mlp_set_norm": (_I, [_P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                              ctypes.c_float, ctypes.c_float]),
    "uhs_mlp_predict": (_I, [_P, ctypes.POINTER(ctypes.c_float), _I, ctypes.POINTER(ctypes.c_float)]),
    "uhs_mlp_grad": (_I, [_P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _I,
                          ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "uhs_mlp_train": (_I, [_P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _I, _I, _I,
                           ctypes.c_float, _I, ctypes.c_uint64, ctypes.POINTER(ctypes.c_float)]),
    "uhs_mlp_adam_step": (_I, [_P, ctypes.POINTER(ctypes.c_float), ctypes.c_float, _I]),
    "uhs_mlp_save": (_I, [_P, ctypes.c_char_p]),
    "uhs_mlp_load": (_I, [_P, ctypes.c_char_p]),
    "uhs_copy_f64": (_I, [_P, _P, ctypes.c_size_t, _P]),
    "uhs_fma_f64": (_I, [_P, ctypes.c_size_t, ctypes.c_double, ctypes.c_double, _I, _P]),
    "uhs_last_error": (ctypes.c_char_p, []),
    "uhs_version": (ctypes.c_char_p, []),
    "uhs_device_count": (_I, []),
    "uhs_device_name": (_I, [ctypes.c_char_p, ctypes.c_size_t]),
    "uhs_malloc": (_I, [ctypes.POINTER(_P), ctypes.c_size_t]),
    "uhs_free": (_I, [_P]),
    "uhs_memcpy_h2d": (_I, [_P, _P, ctypes.c_size_t]),
    "uhs_memcpy_d2h": (_I, [_P, _P, ctypes.c_size_t]),
    "uhs_memset": (_I, [_P, _I, ctypes.c_size_t]),
    "uhs_synchronize": (_I, []),
    "uhs_stream_create": (_I, [ctypes.POINTER(_P)]),
    "uhs_stream_destroy": (_I, [_P]),
    "uhs_stream_synchronize": (_I, [_P]),
    "uhs_event_create": (_I, [ctypes.POINTER(_P)]),
    "uhs_event_destroy": (_I, [_P]),
    "uhs_event_record": (_I, [_P, _P]),
    "uhs_event_elapsed_ms": (_I, [_P, _P, ctypes.POINTER(ctypes.c_float)]),
}

_lib = None


def lib():
    """The loaded libuhs.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libuhs.so not built: run __graft_entry__.build() (or make -C "
                               "uxxxxxxxxh_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise UhsError(rc, lib().uhs_last_error().decode())
    return rc


def ext_array(ext):
    return (ctypes.c_int64 * len(ext))(*[int(e) for e in ext])


class StencilSpec:
    """A stencil (term offsets and coefficients in term order) kept alive for ctypes."""

    def __init__(self, dims, offsets, coeffs, name=None, radius=None):
        self.dims = int(dims)
        self.offsets = np.ascontiguousarray(np.asarray(offsets, dtype=np.int32).reshape(-1, dims))
        self.coeffs = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64).reshape(-1))
        assert len(self.offsets) == len(self.coeffs)
        self.name = name
        self.radius = int(np.abs(self.offsets).max()) if radius is None else radius
        self.c = Stencil(self.dims, len(self.coeffs),
                         self.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                         self.coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    @property
    def npoints(self):
        return len(self.coeffs)

    def with_coeffs(self, coeffs):
        return StencilSpec(self.dims, self.offsets, coeffs, self.name, self.radius)


def builtin_names():
    out = []
    for i in range(check(lib().uhs_builtin_count())):
        n = ctypes.c_char_p()
        check(lib().uhs_builtin_info(i, ctypes.byref(n), None, None, None, None, None, None))
        out.append(n.value.decode())
    return out


def builtin(case, coeffs=None):
    """The built-in benchmark stencil `case` ("3d1r_star", ...), DSL order."""
    L = lib()
    i = check(L.uhs_builtin_find(case.encode()))
    name, sym = ctypes.c_char_p(), ctypes.c_char_p()
    dims, rad, npts = _I(), _I(), _I()
    offp, coefp = ctypes.POINTER(_I)(), ctypes.POINTER(ctypes.c_double)()
    check(L.uhs_builtin_info(i, ctypes.byref(name), ctypes.byref(sym), ctypes.byref(dims), ctypes.byref(rad),
                             ctypes.byref(npts), ctypes.byref(offp), ctypes.byref(coefp)))
    offs = np.ctypeslib.as_array(offp, shape=(npts.value * dims.value,)).copy()
    c = np.ctypeslib.as_array(coefp, shape=(npts.value,)).copy() if coeffs is None else coeffs
    s = StencilSpec(dims.value, offs, c, name.value.decode(), rad.value)
    s.symbol = sym.value.decode()
    return s


class ProgramSpec:
    """A fused program (uhs_program) kept alive for ctypes.  producers are
    (offsets, coeffs) over the input grid, or (offsets, coeffs, sources)
    whose term i reads the input (-1) or an earlier producer; the output
    kernel's term j reads the input (sources[j] == -1) or producer
    sources[j]; cells [lo, hi) are written (halo-inclusive coordinates, axis
    0 first)."""

    def __init__(self, dims, producers, offsets, coeffs, sources, lo, hi):
        self.dims = int(dims)
        self.producers = [StencilSpec(dims, pr[0], pr[1]) for pr in producers]
        self.producer_sources = [np.asarray(pr[2] if len(pr) > 2 else [-1] * len(pr[1]), np.int32)
                                 for pr in producers]
        self.output = StencilSpec(dims, offsets, coeffs)
        self.sources = np.ascontiguousarray(np.asarray(sources, np.int32))
        self.lo, self.hi = [int(v) for v in lo], [int(v) for v in hi]
        self._parr = (Stencil * max(1, len(self.producers)))(*[p.c for p in self.producers])
        psrc = np.full((max(1, len(producers)), MAX_PROGRAM_TERMS), -1, np.int32)
        for k, ps in enumerate(self.producer_sources):
            psrc[k, :len(ps)] = ps
        self._psrc = psrc
        chained = any((ps >= 0).any() for ps in self.producer_sources)
        self.c = Program(self.dims, len(self.producers), self._parr, self.output.c,
                         self.sources.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                         (ctypes.c_int64 * 3)(*(self.lo + [0] * (3 - dims))),
                         (ctypes.c_int64 * 3)(*(self.hi + [0] * (3 - dims))),
                         psrc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if chained else None)

    def producer_tuples(self):
        """[(offsets, coeffs, sources)] for the oracle."""
        return [(p.offsets.tolist(), p.coeffs.tolist(), s.tolist())
                for p, s in zip(self.producers, self.producer_sources)]


PROGRAM_AUTO, PROGRAM_TILED, PROGRAM_POINTWISE = 0, 1, 2


def program_sweep_f64(A, B, shape, prog, timesteps, stream=None, kernel=PROGRAM_AUTO, tile=None):
    if kernel == PROGRAM_AUTO and tile is None:
        check(lib().uhs_program_sweep_f64(A.ptr, B.ptr, ext_array(shape), ctypes.byref(prog.c), timesteps, stream))
        return
    t = (ctypes.c_int * 3)(*tile) if tile is not None else None
    check(lib().uhs_program_sweep_ex_f64(A.ptr, B.ptr, ext_array(shape), ctypes.byref(prog.c), timesteps,
                                         kernel, t, stream))


def program_sweep_f32(A, B, shape, prog, timesteps, compensated=False, stream=None, kernel=PROGRAM_AUTO, tile=None):
    t = (ctypes.c_int * 3)(*tile) if tile is not None else None
    check(lib().uhs_program_sweep_ex_f32(A.ptr, B.ptr, ext_array(shape), ctypes.byref(prog.c), timesteps,
                                         1 if compensated else 0, kernel, t, stream))


def tune_program_f64(A, B, shape, prog, reps=3, stream=None):
    """Times the streaming kernel's candidate tiles on A, B (overwritten):
    (best tile, ms per sweep, tiles timed)."""
    tile = (ctypes.c_int * 3)()
    ms = ctypes.c_double()
    n = ctypes.c_int()
    check(lib().uhs_tune_program_f64(A.ptr, B.ptr, ext_array(shape), ctypes.byref(prog.c), reps, tile,
                                     ctypes.byref(ms), ctypes.byref(n), stream))
    return list(tile), ms.value, n.value


def parse_dsl_program(text=None, path=None):
    """Read a multi-kernel DSL program into (ProgramSpec, DslProgram)."""
    info = DslProgram()
    if path is not None:
        check(lib().uhs_parse_dsl_program_file(path.encode(), ctypes.byref(info)))
    else:
        check(lib().uhs_parse_dsl_program(text.encode(), ctypes.byref(info)))
    d = info.dims
    prods = []
    for k in range(info.nproducers):
        n = info.producer_npoints[k]
        prods.append((np.array(info.producer_offsets[k][:n * d], np.int32), np.array(info.producer_coeffs[k][:n]),
                      np.array(info.producer_sources[k][:n], np.int32)))
    n = info.npoints
    prog = ProgramSpec(d, prods, np.array(info.offsets[:n * d], np.int32), np.array(info.coeffs[:n]),
                       np.array(info.sources[:n], np.int32), list(info.domain_lo[:d]), list(info.domain_hi[:d]))
    return prog, info


def parse_dsl(text=None, path=None):
    """Read UHStencil DSL text (or a .dsl file) into (StencilSpec, DslInfo)."""
    info = DslInfo()
    if path is not None:
        check(lib().uhs_parse_dsl_file(path.encode(), ctypes.byref(info)))
    else:
        check(lib().uhs_parse_dsl(text.encode(), ctypes.byref(info)))
    d, n = info.dims, info.npoints
    offs = np.array(info.offsets[:n * d], dtype=np.int32)
    st = StencilSpec(d, offs, np.array(info.coeffs[:n]), info.name.decode())
    return st, info


def make_coeffs(stencil, seed):
    out = np.zeros(stencil.npoints)
    check(lib().uhs_make_coeffs(stencil.dims, stencil.npoints,
                                stencil.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                seed, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def grid_shape(ext, halo):
    return tuple(int(e) + 2 * halo for e in ext)


class DeviceArray:
    """A device allocation with a numpy-like shape/dtype, freed on close()."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().uhs_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().uhs_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes))

    def download(self):
        a = np.empty(self.shape, self.dtype)
        check(lib().uhs_memcpy_d2h(a.ctypes.data, self.ptr, self.nbytes))
        return a

    def close(self):
        if self.ptr:
            lib().uhs_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _cfg(cfg):
    if cfg is None:
        return None
    if isinstance(cfg, Config):
        return ctypes.byref(cfg)
    return ctypes.byref(Config(*cfg))


def _check_grids(ext, halo, dtype, *arrays):
    """Every device grid must be the halo-inclusive shape of ext and halo, in
    the precision the call computes in: the kernels index the grid from those
    extents, so a smaller allocation would be read and written out of
    bounds."""
    want = grid_shape(ext, halo)
    for a in arrays:
        if tuple(a.shape) != want:
            raise ValueError("grid shape %s does not match extents %s with halo %d (expected %s)"
                             % (tuple(a.shape), list(ext), halo, want))
        if np.dtype(a.dtype) != np.dtype(dtype):
            raise ValueError("grid dtype %s, expected %s" % (np.dtype(a.dtype), np.dtype(dtype)))


def sweep_f64(A, B, ext, halo, stencil, timesteps, cfg=None, stream=None):
    _check_grids(ext, halo, np.float64, A, B)
    check(lib().uhs_sweep_f64(A.ptr, B.ptr, len(ext), ext_array(ext), halo, ctypes.byref(stencil.c),
                              timesteps, _cfg(cfg), stream))


def sweep_f32(A, B, ext, halo, stencil, timesteps, compensated=False, cfg=None, stream=None):
    _check_grids(ext, halo, np.float32, A, B)
    check(lib().uhs_sweep_f32(A.ptr, B.ptr, len(ext), ext_array(ext), halo, ctypes.byref(stencil.c),
                              timesteps, 1 if compensated else 0, _cfg(cfg), stream))


def plan_kind(ext, halo, stencil, dtype=DT_F64, cfg=None):
    buf = ctypes.create_string_buffer(32)
    check(lib().uhs_plan_kind(len(ext), ext_array(ext), halo, ctypes.byref(stencil.c), dtype, _cfg(cfg),
                              buf, 32))
    return buf.value.decode()


def default_config(ext, halo, stencil, dtype=DT_F64):
    c = Config()
    check(lib().uhs_default_config(len(ext), ext_array(ext), halo, ctypes.byref(stencil.c), dtype,
                                   ctypes.byref(c)))
    return c


def one_step_config(ext, halo, stencil, dtype=DT_F64):
    """The default configuration with one launch per timestep."""
    c = Config()
    check(lib().uhs_one_step_config(len(ext), ext_array(ext), halo, ctypes.byref(stencil.c), dtype,
                                    ctypes.byref(c)))
    return c


def fill_uniform(arr, ext, halo, seed, stream=None):
    if np.dtype(arr.dtype) not in (np.dtype(np.float64), np.dtype(np.float32)):
        raise ValueError("fill_uniform takes float64 or float32 grids")
    _check_grids(ext, halo, arr.dtype, arr)
    f = lib().uhs_fill_uniform_f64 if arr.dtype == np.float64 else lib().uhs_fill_uniform_f32
    check(f(arr.ptr, len(ext), ext_array(ext), halo, seed, stream))


def error_norms(a, b, ext, halo, stream=None):
    if np.dtype(a.dtype) not in (np.dtype(np.float64), np.dtype(np.float32)):
        raise ValueError("error_norms compares a float64 or float32 grid with a float64 one")
    _check_grids(ext, halo, a.dtype, a)
    _check_grids(ext, halo, np.float64, b)
    out = (ctypes.c_double * 3)()
    f = lib().uhs_error_norms_f32_f64 if a.dtype == np.float32 else lib().uhs_error_norms_f64
    check(f(a.ptr, b.ptr, len(ext), ext_array(ext), halo, out, stream))
    return tuple(out)


def synchronize():
    check(lib().uhs_synchronize())


def set_lds_poison(on):
    """Debug mode: kernels fill each tile's LDS span with NaN before staging."""
    return check(lib().uhs_set_lds_poison(1 if on else 0))


class Timer:
    """hipEvent pair on one stream."""

    def __init__(self, stream=None):
        self.stream = stream
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().uhs_event_create(ctypes.byref(a)))
        check(lib().uhs_event_create(ctypes.byref(b)))
        self.a, self.b = a, b

    def start(self):
        check(lib().uhs_event_record(self.a, self.stream))

    def stop(self):
        check(lib().uhs_event_record(self.b, self.stream))

    def ms(self):
        v = ctypes.c_float()
        check(lib().uhs_event_elapsed_ms(self.a, self.b, ctypes.byref(v)))
        return v.value


class Batch:
    """A planned, graph-captured batch of 3D FP64 grids (uhs_batch_*)."""

    def __init__(self, grids, timesteps, cfg=None):
        """grids: list of (A, B, ext, halo, StencilSpec) with DeviceArray A/B."""
        self._keep = [g[4] for g in grids]
        arr = (GridDesc * len(grids))()
        for i, (A, B, ext, halo, st) in enumerate(grids):
            arr[i] = GridDesc(A.ptr.value, B.ptr.value, 3, (ctypes.c_int64 * 3)(*ext), halo, st.c)
        self.desc = arr
        self.n = len(grids)
        self.timesteps = timesteps
        h = ctypes.c_void_p()
        check(lib().uhs_batch_create(arr, len(grids), timesteps, ctypes.byref(cfg) if cfg else None,
                                     ctypes.byref(h)))
        self.h = h

    def run(self, stream=None):
        check(lib().uhs_batch_run(self.h, stream))

    def info(self):
        nl, wg = ctypes.c_int(), ctypes.c_int()
        nt, lds = ctypes.c_longlong(), ctypes.c_size_t()
        check(lib().uhs_batch_info(self.h, ctypes.byref(nl), ctypes.byref(nt), ctypes.byref(wg), ctypes.byref(lds)))
        return {"launches": nl.value, "tasks": nt.value, "workgroups": wg.value, "lds": lds.value}

    def tune(self, timesteps=None, stream=None):
        """Times the batch-configuration space; returns (best, best_ms, [(cfg, ms)])."""
        cap = 1024
        cands = (BatchConfig * cap)()
        ms = (ctypes.c_double * cap)()
        best, bms, nc = BatchConfig(), ctypes.c_double(), ctypes.c_int()
        check(lib().uhs_batch_tune(self.desc, self.n, timesteps or self.timesteps, ctypes.byref(best),
                                   ctypes.byref(bms), cands, ms, cap, ctypes.byref(nc), stream))
        return best, bms.value, [(cands[i], ms[i]) for i in range(min(cap, nc.value))]

    def run_each(self, stream=None):
        """The unbatched baseline on the same grids."""
        check(lib().uhs_sweep_each_f64(self.desc, self.n, self.timesteps, stream))

    def close(self):
        if self.h:
            lib().uhs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _grid_descs(grids):
    """GridDesc array for (A, B, ext, halo, StencilSpec) or (ext, halo,
    StencilSpec) tuples (the latter carry no device memory: only for calls
    that touch none)."""
    arr = (GridDesc * len(grids))()
    for i, g in enumerate(grids):
        ext, halo, st = g[-3], g[-2], g[-1]
        A = g[0].ptr.value if len(g) == 5 else None
        B = g[1].ptr.value if len(g) == 5 else None
        arr[i] = GridDesc(A, B, 3, (ctypes.c_int64 * 3)(*ext), halo, st.c)
    return arr


def batch_features(grids, cfg):
    """Cost-model features of (batch, configuration) (uhs_batch_features)."""
    out = np.zeros(20, np.float32)
    check(lib().uhs_batch_features(_grid_descs(grids), len(grids), ctypes.byref(cfg), _fp(out)))
    return out


def batch_tune_list(grids, timesteps, cands, stream=None):
    """Times batch configurations on (A, B, ext, halo, StencilSpec) grids
    (contents overwritten); returns (best, best_ms, [ms per candidate])."""
    arr = (BatchConfig * len(cands))(*cands)
    best, bms = BatchConfig(), ctypes.c_double()
    ms = (ctypes.c_double * len(cands))()
    check(lib().uhs_batch_tune_list(_grid_descs(grids), len(grids), timesteps, arr, len(cands),
                                    ctypes.byref(best), ctypes.byref(bms), ms, stream))
    return best, bms.value, list(ms)


def batch_tune_model(grids, timesteps, model, topk=5, stream=None):
    """Model-pruned batch search; returns (best, best_ms, seconds)."""
    best, bms, sec = BatchConfig(), ctypes.c_double(), ctypes.c_double()
    check(lib().uhs_batch_tune_model(_grid_descs(grids), len(grids), timesteps, model.h, topk, ctypes.byref(best),
                                     ctypes.byref(bms), ctypes.byref(sec), stream))
    return best, bms.value, sec.value


def batch_partition(grids, ngpus, cost=None):
    """Longest-processing-time-first split of a batch over ngpus GPUs
    (uhs_batch_partition).  grids: (ext, halo, StencilSpec) or (A, B, ext,
    halo, StencilSpec) tuples (no device memory is touched); cost: per-grid
    costs, or None to estimate them.  Returns (gpu of each grid, load of each
    GPU)."""
    n = len(grids)
    arr = (GridDesc * n)()
    for i, g in enumerate(grids):
        ext, halo, st = g[-3], g[-2], g[-1]
        arr[i] = GridDesc(None, None, 3, (ctypes.c_int64 * 3)(*ext), halo, st.c)
    gpu = (ctypes.c_int * n)()
    load = (ctypes.c_double * max(1, ngpus))()
    c = (ctypes.c_double * n)(*cost) if cost is not None else None
    check(lib().uhs_batch_partition(arr, n, ngpus, c, gpu, load))
    return list(gpu), list(load)[:ngpus]


def batch_space():
    n = check(lib().uhs_batch_space(None, 0))
    arr = (BatchConfig * n)()
    check(lib().uhs_batch_space(arr, n))
    return [arr[i] for i in range(n)]


def batch_default_config():
    c = BatchConfig()
    check(lib().uhs_batch_default_config(ctypes.byref(c)))
    return c


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def tune_space(ext, halo, stencil, dtype=DT_F64):
    L = lib()
    n = check(L.uhs_tune_space(len(ext), ext_array(ext), halo, ctypes.byref(stencil.c), dtype, None, 0))
    arr = (Config * max(1, n))()
    check(L.uhs_tune_space(len(ext), ext_array(ext), halo, ctypes.byref(stencil.c), dtype, arr, n))
    return [Config(*arr[i].tuple()) for i in range(n)]


def tune_sweep(A, B, ext, halo, stencil, dtype=DT_F64, reps=3, db=None, model=None, topk=5, stream=None):
    _check_grids(ext, halo, np.float64 if dtype == DT_F64 else np.float32, A, B)
    r = TuneResult()
    dbp = db.encode() if db else None
    if model is None:
        check(lib().uhs_tune_sweep(A.ptr, B.ptr, len(ext), ext_array(ext), halo, ctypes.byref(stencil.c), dtype,
                                   reps, dbp, ctypes.byref(r), stream))
    elif isinstance(model, (list, tuple)):
        # An ensemble: configurations ranked by the mean of the members'
        # predicted log times.
        hs = (_P * len(model))(*[m.h for m in model])
        check(lib().uhs_tune_sweep_models(A.ptr, B.ptr, len(ext), ext_array(ext), halo, ctypes.byref(stencil.c),
                                          dtype, hs, len(model), topk, reps, dbp, ctypes.byref(r), stream))
    else:
        check(lib().uhs_tune_sweep_model(A.ptr, B.ptr, len(ext), ext_array(ext), halo, ctypes.byref(stencil.c),
                                         dtype, model.h, topk, reps, dbp, ctypes.byref(r), stream))
    return r


# uhs_tune_timer_fn: int (*)(const uhs_config *, int reps, double *ms, void *user)
TimerFn = ctypes.CFUNCTYPE(_I, ctypes.POINTER(Config), _I, ctypes.POINTER(ctypes.c_double), _P)
_timer_ref = None


def tune_set_timer(fn):
    """Mock timing backend for the searches: fn(cfg, reps) -> (status, ms)
    replaces the timed sweeps (None restores the GPU timer).  The search
    then runs without a GPU and never touches its grids."""
    global _timer_ref
    if fn is None:
        check(lib().uhs_tune_set_timer(None, None))
        _timer_ref = None
        return

    def cb(cfg, reps, ms, user):
        status, t = fn(Config(cfg[0].tile_x, cfg[0].tile_y, cfg[0].tile_z, cfg[0].vec, cfg[0].flags), reps)
        ms[0] = t
        return status

    ref = TimerFn(cb)
    check(lib().uhs_tune_set_timer(ctypes.cast(ref, _P), None))
    _timer_ref = ref


def tune_list(A, B, ext, halo, stencil, cands, dtype=DT_F64, reps=2, db=None, stream=None):
    _check_grids(ext, halo, np.float64 if dtype == DT_F64 else np.float32, A, B)
    arr = (Config * len(cands))(*[c if isinstance(c, Config) else Config(*c) for c in cands])
    r = TuneResult()
    check(lib().uhs_tune_sweep_list(A.ptr, B.ptr, len(ext), ext_array(ext), halo, ctypes.byref(stencil.c), dtype,
                                    arr, len(cands), reps, db.encode() if db else None, ctypes.byref(r), stream))
    return r


def tune_time(A, B, ext, halo, stencil, cfg, dtype=DT_F64, reps=3, stream=None):
    _check_grids(ext, halo, np.float64 if dtype == DT_F64 else np.float32, A, B)
    ms = ctypes.c_double()
    c = cfg if isinstance(cfg, Config) else Config(*cfg)
    check(lib().uhs_tune_time(A.ptr, B.ptr, len(ext), ext_array(ext), halo, ctypes.byref(stencil.c), dtype,
                              ctypes.byref(c), reps, ctypes.byref(ms), stream))
    return ms.value


def mlp_features(dims, radius, star, ext, cfg, depth=1, dtype=DT_F64):
    out = np.zeros(20, np.float32)
    c = cfg if isinstance(cfg, Config) else Config(*cfg)
    e = list(ext) + [0] * (3 - len(ext))
    check(lib().uhs_mlp_features(dims, radius, 1 if star else 0, dtype, ext_array(e), ctypes.byref(c), depth,
                                 _fp(out)))
    return out


class MLP:
    """The cost model (uhs_mlp_*): parameters and training live on the GPU."""

    def __init__(self, seed=1):
        h = ctypes.c_void_p()
        check(lib().uhs_mlp_create(seed, ctypes.byref(h)))
        self.h = h
        self.nparam = lib().uhs_mlp_num_params()

    def params(self):
        out = np.zeros(self.nparam, np.float32)
        check(lib().uhs_mlp_get_params(self.h, _fp(out), self.nparam))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(lib().uhs_mlp_set_params(self.h, _fp(p), self.nparam))

    def set_norm(self, mean, scale, tmean, tscale):
        m = np.ascontiguousarray(mean, np.float32)
        s = np.ascontiguousarray(scale, np.float32)
        check(lib().uhs_mlp_set_norm(self.h, _fp(m), _fp(s), tmean, tscale))

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(len(x), np.float32)
        check(lib().uhs_mlp_predict(self.h, _fp(x), len(x), _fp(y)))
        return y

    def grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.nparam, np.float32)
        loss = ctypes.c_float()
        check(lib().uhs_mlp_grad(self.h, _fp(x), _fp(t), len(x), _fp(g), ctypes.byref(loss)))
        return g, loss.value

    def adam_step(self, g, lr, t):
        g = np.ascontiguousarray(g, np.float32)
        check(lib().uhs_mlp_adam_step(self.h, _fp(g), lr, t))

    def train(self, x, t, epochs=100, batch=256, lr=1e-3, fit_norm=True, seed=1):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = ctypes.c_float()
        check(lib().uhs_mlp_train(self.h, _fp(x), _fp(t), len(x), epochs, batch, lr, 1 if fit_norm else 0, seed,
                                  ctypes.byref(loss)))
        return loss.value

    def save(self, path):
        check(lib().uhs_mlp_save(self.h, path.encode()))

    def load(self, path):
        check(lib().uhs_mlp_load(self.h, path.encode()))

    def close(self):
        if self.h:
            lib().uhs_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
