"""ctypes view of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "liboracle.so")
PLAIN, FUSED, KAHAN = 0, 1, 2

_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


class Oracle:
    def __init__(self, path=PATH):
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: make -C oracle")
        L = ctypes.CDLL(path)
        I = ctypes.c_int
        L.orc_sweep_f64.argtypes = [_dp, _dp, I, _i64, I, I, _ip, _dp, I, I]
        L.orc_sweep_f32.argtypes = [_fp, _fp, I, _i64, I, I, _ip, _fp, I, I]
        L.orc_iterate_f64.argtypes = [_dp, _dp, I, _i64, I, I, _ip, _dp, I, I, I]
        L.orc_iterate_f32.argtypes = [_fp, _fp, I, _i64, I, I, _ip, _fp, I, I, I]
        L.orc_fill_uniform_f64.argtypes = [_dp, I, _i64, I, ctypes.c_uint64]
        L.orc_fill_uniform_f32.argtypes = [_fp, I, _i64, I, ctypes.c_uint64]
        L.orc_make_coeffs.argtypes = [I, I, _ip, ctypes.c_uint64, _dp]
        L.orc_fill_rand50_f64.argtypes = [_dp, ctypes.c_size_t]
        L.orc_program_f64.argtypes = [_dp, _dp, I, _i64, I, _ip, _ip, _dp, _ip, I, _ip, _dp, _ip, _i64, _i64, I, I]
        L.orc_program_f64.restype = I
        L.orc_program_f32.argtypes = [_fp, _fp, I, _i64, I, _ip, _ip, _fp, _ip, I, _ip, _fp, _ip, _i64, _i64, I, I]
        L.orc_program_f32.restype = I
        L.orc_mix64.argtypes = [ctypes.c_uint64]
        L.orc_mix64.restype = ctypes.c_uint64
        for f in ("orc_sweep_f64", "orc_sweep_f32", "orc_iterate_f64", "orc_iterate_f32"):
            getattr(L, f).restype = I
        self.L = L

    @staticmethod
    def _st(offsets, dims):
        return np.ascontiguousarray(np.asarray(offsets, np.int32).reshape(-1, dims))

    def sweep(self, a, b, ext, halo, offsets, coeffs, mode=FUSED, nthreads=0):
        """One sweep a -> b (b modified in place)."""
        dims = len(ext)
        off = self._st(offsets, dims)
        e = np.asarray(ext, np.int64)
        if a.dtype == np.float64:
            rc = self.L.orc_sweep_f64(a, b, dims, e, halo, len(off), off,
                                      np.asarray(coeffs, np.float64), mode, nthreads)
        else:
            rc = self.L.orc_sweep_f32(a, b, dims, e, halo, len(off), off,
                                      np.asarray(coeffs, np.float64).astype(np.float32), mode, nthreads)
        assert rc == 0

    def iterate(self, A, B, ext, halo, offsets, coeffs, T, mode=FUSED, nthreads=0):
        dims = len(ext)
        off = self._st(offsets, dims)
        e = np.asarray(ext, np.int64)
        if A.dtype == np.float64:
            rc = self.L.orc_iterate_f64(A, B, dims, e, halo, len(off), off,
                                        np.asarray(coeffs, np.float64), T, mode, nthreads)
        else:
            rc = self.L.orc_iterate_f32(A, B, dims, e, halo, len(off), off,
                                        np.asarray(coeffs, np.float64).astype(np.float32), T, mode, nthreads)
        assert rc == 0

    def program(self, A, B, shape, producers, offsets, coeffs, sources, lo, hi, T, mode=FUSED):
        """T sweeps of a fused program (A, B modified in place); producers =
        [(offsets, coeffs) or (offsets, coeffs, sources), ...]."""
        dims = len(shape)
        pnp = np.array([len(np.asarray(pr[1])) for pr in producers] or [0], np.int32)
        poff = np.ascontiguousarray(np.concatenate([self._st(pr[0], dims).reshape(-1) for pr in producers])
                                    if producers else np.zeros(1, np.int32), dtype=np.int32)
        pcoef = np.ascontiguousarray(np.concatenate([np.asarray(pr[1], np.float64) for pr in producers])
                                     if producers else np.zeros(1))
        psrc = np.ascontiguousarray(np.concatenate(
            [np.asarray(pr[2] if len(pr) > 2 else [-1] * len(pr[1]), np.int32) for pr in producers])
            if producers else np.full(1, -1, np.int32), dtype=np.int32)
        off = self._st(offsets, dims)
        if A.dtype == np.float32:   # FP32 storage: coefficients rounded to float
            return self.L.orc_program_f32(A, B, dims, np.asarray(shape, np.int64), len(producers), pnp, poff,
                                          pcoef.astype(np.float32), psrc, len(off), off,
                                          np.asarray(coeffs, np.float64).astype(np.float32),
                                          np.asarray(sources, np.int32), np.asarray(lo, np.int64),
                                          np.asarray(hi, np.int64), T, mode)
        rc = self.L.orc_program_f64(A, B, dims, np.asarray(shape, np.int64), len(producers), pnp, poff, pcoef,
                                    psrc, len(off), off, np.asarray(coeffs, np.float64),
                                    np.asarray(sources, np.int32), np.asarray(lo, np.int64),
                                    np.asarray(hi, np.int64), T, mode)
        return rc

    def fill(self, ext, halo, seed, dtype=np.float64):
        a = np.empty([int(e) + 2 * halo for e in ext], dtype)
        f = self.L.orc_fill_uniform_f64 if dtype == np.float64 else self.L.orc_fill_uniform_f32
        f(a, len(ext), np.asarray(ext, np.int64), halo, seed)
        return a

    def make_coeffs(self, offsets, dims, seed):
        off = self._st(offsets, dims)
        c = np.zeros(len(off))
        self.L.orc_make_coeffs(dims, len(off), off, seed, c)
        return c

    def mix64(self, z):
        return self.L.orc_mix64(z)
