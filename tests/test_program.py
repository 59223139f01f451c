"""Fused two-level programs on the CPU: the oracle's nested evaluation
(orc_program_f64) against an independent numpy two-pass evaluation, and the
multi-kernel DSL reader (uhs_parse_dsl_program).

The two-pass evaluation is the reference's own definition of a multi-kernel
program: the producer is written to a temporary grid, then the consumer reads
it (UHStencil/examples/2d5pt_nested/stencil_2d5pt_nested_driver.serial.c:40-60,
the driver's check loop; stencil_2d5pt_nested_kernel.serial_master.c:18-32,
the generated code's tmp buffer).  Plain mode evaluates each kernel as a
left-associated sum of separate products, like that C code, so it must match
the two-pass evaluation bitwise; fused mode (the GPU kernel's fma chain) is
held to 1e-12 of the sum of the terms' magnitudes."""
import os

import numpy as np
import pytest

from tests.oracle_py import FUSED, KAHAN, PLAIN

REF = "/root/reference/UHStencil/examples"


def slab(lo, hi, off):
    return tuple(slice(l + o, h + o) for l, h, o in zip(lo, hi, off))


MARGIN = 12


def materialise(inp, producers, absolute=False):
    """Every producer written to a temporary grid, in dependency order (the
    reference's unfused lowering).  The temporaries extend MARGIN cells past
    the input grid on every side: a fused producer may be evaluated at a
    cell outside the grid whose own reads are inside.  Cells whose reads
    leave the grid hold NaN.  absolute: the same sums of |terms|."""
    dims, M = inp.ndim, MARGIN
    dt = inp.dtype.type                      # the arithmetic type: float64 or float32
    src_in = np.abs(inp) if absolute else inp
    P, reach = [], []
    for pr in producers:
        offs, coefs = pr[0], pr[1]
        srcs = pr[2] if len(pr) > 2 else [-1] * len(coefs)
        rlo = [min(w[a] + (reach[s][0][a] if s >= 0 else 0) for w, s in zip(offs, srcs)) for a in range(dims)]
        rhi = [max(w[a] + (reach[s][1][a] if s >= 0 else 0) for w, s in zip(offs, srcs)) for a in range(dims)]
        reach.append((rlo, rhi))
        # Valid cells in grid coordinates, limited to the margin.
        lo = [max(-M, -r) for r in rlo]
        hi = [min(n + M, n - r) for n, r in zip(inp.shape, rhi)]
        arr = np.full([n + 2 * M for n in inp.shape], np.nan, dtype=inp.dtype)
        if all(h > l for l, h in zip(lo, hi)):
            acc = None
            for w, c, s in zip(offs, coefs, srcs):
                if s < 0:
                    v = src_in[slab(lo, hi, w)]
                else:
                    v = P[s][slab([l + M for l in lo], [h + M for h in hi], w)]
                t = dt(abs(c) if absolute else c) * v
                acc = t if acc is None else acc + t
            arr[slab([l + M for l in lo], [h + M for h in hi], [0] * dims)] = acc
        P.append(arr)
    return P


def two_pass(inp, producers, offsets, coeffs, sources, lo, hi, absolute=False):
    """One sweep over [lo, hi) the unfused way: producers materialised as
    temporary grids, then the output kernel reads them."""
    P = materialise(inp, producers, absolute)
    M = MARGIN
    dt = inp.dtype.type
    out = None
    for v, c, s in zip(offsets, coeffs, sources):
        if s < 0:
            x = (np.abs(inp) if absolute else inp)[slab(lo, hi, v)]
        else:
            x = P[s][slab([l + M for l in lo], [h + M for h in hi], v)]
        t = dt(abs(c) if absolute else c) * x
        out = t if out is None else out + t
    assert not np.isnan(out).any()
    return out


def random_program(rng, dims, chains=True):
    """A random program: 0-4 producers; with chains, a producer's terms read
    the input or earlier producers (offsets in [-1, 1] then, so the reach
    stays small)."""
    nprod = int(rng.integers(0, 5))
    prods = []
    for k in range(nprod):
        n = int(rng.integers(1, 7))
        chained = chains and k > 0 and rng.random() < 0.6
        srcs = [int(rng.integers(-1, k)) if chained and rng.random() < 0.7 else -1 for _ in range(n)]
        span = 2 if chained else 3
        prods.append((rng.integers(-span + 1, span, size=(n, dims)).tolist(), rng.uniform(-1, 1, n).tolist(), srcs))
    n = int(rng.integers(1, 8))
    offs = rng.integers(-2, 3, size=(n, dims)).tolist()
    coefs = rng.uniform(-1, 1, n).tolist()
    srcs = rng.integers(-1, nprod, size=n).tolist() if nprod else [-1] * n
    shape = [int(rng.integers(14, 22)) for _ in range(dims)] if dims == 3 else \
        [int(rng.integers(24, 44)), int(rng.integers(24, 44))]
    # Reach of every term (through producer chains), for a domain whose
    # reads stay inside the grid.
    preach = []
    for po, pc, ps in prods:
        lo_ = [min(w[a] + (preach[s][0][a] if s >= 0 else 0) for w, s in zip(po, ps)) for a in range(dims)]
        hi_ = [max(w[a] + (preach[s][1][a] if s >= 0 else 0) for w, s in zip(po, ps)) for a in range(dims)]
        preach.append((lo_, hi_))
    rlo, rhi = [0] * dims, [0] * dims
    for j in range(n):
        for a in range(dims):
            s = srcs[j]
            rlo[a] = min(rlo[a], offs[j][a] + (preach[s][0][a] if s >= 0 else 0))
            rhi[a] = max(rhi[a], offs[j][a] + (preach[s][1][a] if s >= 0 else 0))
    lo = [-rlo[a] + int(rng.integers(0, 3)) for a in range(dims)]
    hi = [shape[a] - rhi[a] - int(rng.integers(0, 3)) for a in range(dims)]
    if any(h <= l for l, h in zip(lo, hi)):
        return random_program(rng, dims, chains)
    return shape, prods, offs, coefs, srcs, lo, hi


@pytest.mark.parametrize("seed", range(12))
def test_oracle_plain_is_two_pass(orc, seed):
    rng = np.random.default_rng(seed)
    dims = 2 + seed % 2
    shape, prods, offs, coefs, srcs, lo, hi = random_program(rng, dims)
    A0, B0 = rng.random(shape), rng.random(shape)
    for T in (1, 2, 3):
        A, B = A0.copy(), B0.copy()
        assert orc.program(A, B, shape, prods, offs, coefs, srcs, lo, hi, T, PLAIN) == 0
        eA, eB = A0.copy(), B0.copy()
        for s in range(1, T + 1):
            src, dst = (eA, eB) if s & 1 else (eB, eA)
            dst[slab(lo, hi, [0] * dims)] = two_pass(src, prods, offs, coefs, srcs, lo, hi)
        assert np.array_equal(A, eA) and np.array_equal(B, eB), (seed, T)


def max_program(rng, dims):
    """The largest program the interface takes: 4 producers of 32 terms and
    an output kernel of 32 terms, 8 from each producer (offsets in [-1, 1])."""
    prods = [(rng.integers(-1, 2, size=(32, dims)).tolist(), rng.uniform(-1, 1, 32).tolist(), [-1] * 32)
             for _ in range(4)]
    offs = rng.integers(-1, 2, size=(32, dims)).tolist()
    coefs = rng.uniform(-1, 1, 32).tolist()
    srcs = [j % 4 for j in range(32)]
    shape = [14, 15, 70] if dims == 3 else [40, 150]
    return shape, prods, offs, coefs, srcs, [2] * dims, [n - 2 for n in shape]


@pytest.mark.parametrize("dims", [2, 3])
def test_oracle_chain_program(orc, dims):
    from tests.test_gpu_program import chain_program
    shape, prods, offs, coefs, srcs, lo, hi = chain_program(dims)
    rng = np.random.default_rng(5 + dims)
    A0, B0 = rng.random(shape), rng.random(shape)
    A, B = A0.copy(), B0.copy()
    assert orc.program(A, B, shape, prods, offs, coefs, srcs, lo, hi, 1, PLAIN) == 0
    assert np.array_equal(B[slab(lo, hi, [0] * dims)], two_pass(A0, prods, offs, coefs, srcs, lo, hi))
    # A producer reading a later one is refused.
    bad = [prods[0], (prods[1][0], prods[1][1], [2, -1, 0])] + prods[2:]
    assert orc.program(A, B, shape, bad, offs, coefs, srcs, lo, hi, 1, PLAIN) == -1


@pytest.mark.parametrize("dims", [2, 3])
def test_oracle_max_program(orc, dims):
    rng = np.random.default_rng(40 + dims)
    shape, prods, offs, coefs, srcs, lo, hi = max_program(rng, dims)
    A0, B0 = rng.random(shape), rng.random(shape)
    A, B = A0.copy(), B0.copy()
    assert orc.program(A, B, shape, prods, offs, coefs, srcs, lo, hi, 1, PLAIN) == 0
    want = two_pass(A0, prods, offs, coefs, srcs, lo, hi)
    assert np.array_equal(B[slab(lo, hi, [0] * dims)], want)


@pytest.mark.parametrize("seed", range(8))
def test_oracle_plain_f32_is_two_pass(orc, seed):
    """FP32: the oracle's plain mode is the float32 two-pass evaluation
    (coefficients rounded to float), chains included."""
    rng = np.random.default_rng(200 + seed)
    dims = 2 + seed % 2
    shape, prods, offs, coefs, srcs, lo, hi = random_program(rng, dims)
    A0 = rng.random(shape).astype(np.float32)
    B0 = rng.random(shape).astype(np.float32)
    A, B = A0.copy(), B0.copy()
    assert orc.program(A, B, shape, prods, offs, coefs, srcs, lo, hi, 1, PLAIN) == 0
    assert B.dtype == np.float32
    assert np.array_equal(B[slab(lo, hi, [0] * dims)], two_pass(A0, prods, offs, coefs, srcs, lo, hi))


def test_oracle_kahan_f32_closed_form(orc):
    """Compensated FP32 sums, closed form: on a field of ones, a producer of
    1 followed by 7 terms of 2^-25 gives exactly 1 in plain FP32 (each
    2^-25 is below half an ulp of 1) and 1 + 7 * 2^-25 compensated; the
    output kernel reads it once with coefficient 1."""
    shape, lo, hi = [12, 20], [2, 2], [10, 18]
    offs = [[0, 0], [0, 1], [1, 0], [0, -1], [-1, 0], [1, 1], [-1, -1], [1, -1]]
    prods = [(offs, [1.0] + [2.0 ** -25] * 7, [-1] * 8)]
    for mode, want in ((PLAIN, np.float32(1.0)), (KAHAN, np.float32(1.0 + 7 * 2.0 ** -25))):
        A = np.ones(shape, np.float32)
        B = np.zeros(shape, np.float32)
        assert orc.program(A, B, shape, prods, [[0, 0]], [1.0], [0], lo, hi, 1, mode) == 0
        assert np.all(B[2:10, 2:18] == want), mode
    # FP64 has no compensated mode.
    A, B = np.ones(shape), np.zeros(shape)
    assert orc.program(A, B, shape, prods, [[0, 0]], [1.0], [0], lo, hi, 1, KAHAN) == -1


@pytest.mark.parametrize("seed", range(6))
def test_oracle_fused_within_bound(orc, seed):
    rng = np.random.default_rng(100 + seed)
    dims = 2 + seed % 2
    shape, prods, offs, coefs, srcs, lo, hi = random_program(rng, dims)
    A0, B0 = rng.random(shape), rng.random(shape)
    A, B = A0.copy(), B0.copy()
    assert orc.program(A, B, shape, prods, offs, coefs, srcs, lo, hi, 1, FUSED) == 0
    want = two_pass(A0, prods, offs, coefs, srcs, lo, hi)
    mag = two_pass(A0, prods, offs, coefs, srcs, lo, hi, absolute=True)
    got = B[slab(lo, hi, [0] * dims)]
    assert np.all(np.abs(got - want) <= 1e-12 * np.maximum(mag, 1e-300))
    # Outside the domain nothing is written.
    mask = np.ones(shape, bool)
    mask[slab(lo, hi, [0] * dims)] = False
    assert np.array_equal(B[mask], B0[mask]) and np.array_equal(A, A0)


def test_oracle_detects_planted_index_error(orc):
    """A consumer that reads its producer at the wrong offset (one cell off)
    must not match: the two-pass check sees the offsets, not just the sums."""
    rng = np.random.default_rng(5)
    shape, lo, hi = [30, 30], [4, 4], [26, 26]
    prods = [([[-1, 0], [0, 1], [0, 0]], [0.2, 0.4, -0.5])]
    offs, coefs, srcs = [[-1, -1], [1, 1], [0, 0]], [0.1, 0.3, -0.5], [0, 0, 0]
    A0, B0 = rng.random(shape), rng.random(shape)
    want = two_pass(A0, prods, offs, coefs, srcs, lo, hi)
    A, B = A0.copy(), B0.copy()
    assert orc.program(A, B, shape, prods, [[-1, -1], [1, 0], [0, 0]], coefs, srcs, lo, hi, 1, PLAIN) == 0
    assert not np.array_equal(B[4:26, 4:26], want)


def test_oracle_rejects_reads_outside_grid(orc):
    A, B = np.zeros((10, 10)), np.zeros((10, 10))
    prods = [([[0, -2]], [1.0])]
    # producer reach -2 plus term offset -1 at lo = 2 reads column -1
    assert orc.program(A, B, [10, 10], prods, [[0, -1]], [1.0], [0], [2, 2], [8, 8], 1, PLAIN) == -1
    assert orc.program(A, B, [10, 10], prods, [[0, -1]], [1.0], [0], [2, 3], [8, 8], 1, PLAIN) == 0
    assert orc.program(A, B, [10, 10], prods, [[0, 0]], [1.0], [1], [2, 3], [8, 8], 1, PLAIN) == -1


NESTED_2D = """
stencil nested_small(double input[98][98]) {
    iteration(4)
    operation (outer)
    kernel inner {
        domain([1, 97][1, 97])
        expr {
            0.2*input[y-1][x]
            + 0.4*input[y][x-1] - 0.5*input[y][x] + 0.4*input[y][x+1]
            + 0.2*input[y+1][x]
        }
    }
    kernel outer {
        domain([17, 81][17, 81])
        expr {
            0.1*inner[y-1][x-1] + 0.3*inner[y-1][x+1]
            - 0.5*inner[y][x]
            + 0.3*inner[y+1][x-1] + 0.1*inner[y+1][x+1]
        }
    }
}
"""


def in_place_check_loop(inp, T, h2):
    """The reference driver's own check loop, restated in numpy: each
    iteration writes tmp over the whole interior, then the consumer back into
    the input over the inner domain (stencil_2d5pt_nested_driver.serial.c:44-60)."""
    a = inp.copy()
    tmp = np.zeros_like(a)
    M, N = a.shape
    for _ in range(T):
        tmp[1:M - 1, 1:N - 1] = (0.2 * a[0:M - 2, 1:N - 1] + 0.4 * a[1:M - 1, 0:N - 2]
                                 - 0.5 * a[1:M - 1, 1:N - 1] + 0.4 * a[1:M - 1, 2:N]
                                 + 0.2 * a[2:M, 1:N - 1])
        lo, hi = 1 + h2, M - 1 - h2
        a[lo:hi, lo:hi] = (0.1 * tmp[lo - 1:hi - 1, lo - 1:hi - 1] + 0.3 * tmp[lo - 1:hi - 1, lo + 1:hi + 1]
                           - 0.5 * tmp[lo:hi, lo:hi]
                           + 0.3 * tmp[lo + 1:hi + 1, lo - 1:hi - 1] + 0.1 * tmp[lo + 1:hi + 1, lo + 1:hi + 1])
    return a


def run_parsed(orc, prog, A, B, T, mode):
    prods = prog.producer_tuples()
    return orc.program(A, B, [int(v) for v in A.shape], prods, prog.output.offsets.tolist(),
                       prog.output.coeffs.tolist(), prog.sources.tolist(), prog.lo, prog.hi, T, mode)


def test_parse_nested_program_and_reference_check_loop(uhs, orc):
    prog, info = uhs.parse_dsl_program(NESTED_2D)
    assert info.name.decode() == "nested_small" and info.operation.decode() == "outer"
    assert info.nproducers == 1 and info.producer_names[0].value.decode() == "inner"
    assert prog.producers[0].offsets.tolist() == [[-1, 0], [0, -1], [0, 0], [0, 1], [1, 0]]
    assert prog.producers[0].coeffs.tolist() == [0.2, 0.4, -0.5, 0.4, 0.2]
    assert prog.output.offsets.tolist() == [[-1, -1], [-1, 1], [0, 0], [1, -1], [1, 1]]
    assert prog.output.coeffs.tolist() == [0.1, 0.3, -0.5, 0.3, 0.1]
    assert prog.sources.tolist() == [0] * 5 and prog.lo == [17, 17] and prog.hi == [81, 81]
    assert info.iteration == 4 and list(info.grid[:2]) == [98, 98]
    # Ping-pong with B a copy of A equals the driver's in-place loop: cells
    # outside the domain are never written, so both grids keep them.
    rng = np.random.default_rng(3)
    A = rng.random((98, 98))
    B = A.copy()
    want = in_place_check_loop(A, 4, 16)
    assert run_parsed(orc, prog, A, B, 4, PLAIN) == 0
    assert np.array_equal(A, want)


def test_default_operation_and_input_terms(uhs):
    text = """stencil s(double a[12][14]) { iteration(2)
      kernel k1 { domain([1,11][1,13]) expr { 0.2*a[y-1][x] + 0.4*a[y][x-1] - 0.5*a[y][x] } }
      kernel k2 { domain([3,9][3,11]) expr { 0.1*k1[y-1][x-1] + 0.3*a[y][x+1] - 0.5*k1[y][x] } } }"""
    prog, info = uhs.parse_dsl_program(text)
    assert info.operation.decode() == "k2" and info.nproducers == 1
    assert prog.sources.tolist() == [0, -1, 0]
    assert prog.output.offsets.tolist() == [[-1, -1], [0, 1], [0, 0]]
    assert prog.lo == [3, 3] and prog.hi == [9, 11]


def test_program_view_points_into_the_parsed_program(uhs):
    """uhs_dsl_program_view: the uhs_program a C caller hands to
    uhs_program_sweep_f64, built from the parsed struct."""
    import ctypes
    _, info = uhs.parse_dsl_program(NESTED_2D)
    prods = (uhs.Stencil * uhs.MAX_PRODUCERS)()
    pg = uhs.Program()
    uhs.check(uhs.lib().uhs_dsl_program_view(ctypes.byref(info), prods, ctypes.byref(pg)))
    assert pg.dims == 2 and pg.nproducers == 1 and list(pg.lo[:2]) == [17, 17] and list(pg.hi[:2]) == [81, 81]
    assert pg.output.npoints == 5 and [pg.output.coeffs[i] for i in range(5)] == [0.1, 0.3, -0.5, 0.3, 0.1]
    assert [pg.sources[i] for i in range(5)] == [0] * 5
    p0 = pg.producers[0]
    assert p0.npoints == 5 and [p0.offsets[i] for i in range(10)] == [-1, 0, 0, -1, 0, 0, 0, 1, 1, 0]
    assert ctypes.addressof(p0.coeffs.contents) == ctypes.addressof(info.producer_coeffs[0])
    with pytest.raises(uhs.UhsError):
        uhs.check(uhs.lib().uhs_dsl_program_view(ctypes.byref(info), None, ctypes.byref(pg)))


def test_single_kernel_program_matches_parse_dsl(uhs):
    text = """stencil one(double f[10][12][14]) { iteration(4)
      kernel k { expr { 0.5*f[z-1][y][x] + 0.25*f[z][y][x+1] - f[z+1][y-1][x] } } }"""
    prog, info = uhs.parse_dsl_program(text)
    st, sinfo = uhs.parse_dsl(text)
    assert info.nproducers == 0 and prog.sources.tolist() == [-1, -1, -1]
    assert prog.output.offsets.tolist() == st.offsets.tolist()
    assert prog.output.coeffs.tolist() == st.coeffs.tolist()
    # Default domain: the cells whose reads stay inside, per axis and side
    # (z reads -1 and +1, y only -1, x only +1).
    assert prog.lo == [1, 1, 0] and prog.hi == [9, 12, 13]


def test_program_reader_chains(uhs, orc):
    """Chains fuse completely: k3 reads k2, which reads k1 and the array."""
    chain = """stencil c(double a[20][24]) { operation(k3)
      kernel k1 { expr { 0.5*a[y][x+1] + 0.25*a[y-1][x] } }
      kernel k2 { expr { 0.5*k1[y][x] - 0.75*a[y][x-1] + 0.125*k1[y+1][x-1] } }
      kernel unused { expr { a[y][x] } }
      kernel k3 { domain([3,17][3,21]) expr { 0.5*k2[y][x-1] + 2.0*a[y][x] + k2[y-1][x+1] } } }"""
    prog, info = uhs.parse_dsl_program(chain)
    assert info.nproducers == 2
    assert [info.producer_names[i].value.decode() for i in range(2)] == ["k1", "k2"]
    assert prog.producer_sources[0].tolist() == [-1, -1]
    assert prog.producer_sources[1].tolist() == [0, -1, 0]
    assert prog.sources.tolist() == [1, -1, 1]
    rng = np.random.default_rng(2)
    A0, B0 = rng.random((20, 24)), rng.random((20, 24))
    A, B = A0.copy(), B0.copy()
    assert run_parsed(orc, prog, A, B, 1, PLAIN) == 0
    assert np.array_equal(B[3:17, 3:21], two_pass(A0, prog.producer_tuples(), prog.output.offsets.tolist(),
                                                   prog.output.coeffs.tolist(), prog.sources.tolist(), [3, 3], [17, 21]))


def test_program_reader_errors(uhs):
    forward = """stencil c(double a[20][20]) { operation(k2)
      kernel k1 { expr { 0.5*k2[y][x] } }
      kernel k2 { expr { 0.5*k1[y][x-1] } } }"""
    with pytest.raises(uhs.UhsError, match="earlier kernel"):
        uhs.parse_dsl_program(forward)
    unknown = """stencil u(double a[20][20]) { kernel k { expr { 0.5*b[y][x] } } }"""
    with pytest.raises(uhs.UhsError):
        uhs.parse_dsl_program(unknown)
    missing_op = """stencil m(double a[20][20]) { operation(nope) kernel k { expr { a[y][x] } } }"""
    with pytest.raises(uhs.UhsError):
        uhs.parse_dsl_program(missing_op)
    five = "stencil f(double a[20][20]) { " + " ".join(
        "kernel p%d { expr { 0.5*a[y][x] } }" % i for i in range(5)) + \
        " kernel o { expr { " + " + ".join("p%d[y][x]" % i for i in range(5)) + " } } }"
    with pytest.raises(uhs.UhsError, match="more than 4"):
        uhs.parse_dsl_program(five)


class HostBuf:
    """Host memory standing in for a device grid: every call below fails its
    argument checks before any launch, so nothing reads it."""

    def __init__(self, shape):
        self.a = np.zeros(shape)
        self.ptr = self.a.ctypes.data


def test_program_sweep_rejects_bad_programs(uhs):
    A, B = HostBuf((16, 16)), HostBuf((16, 16))

    def sweep(prods, offs, coefs, srcs, lo, hi, a=A, b=B, T=1):
        uhs.program_sweep_f64(a, b, [16, 16], uhs.ProgramSpec(2, prods, offs, coefs, srcs, lo, hi), T)

    with pytest.raises(uhs.UhsError, match="outside the grid"):
        sweep([([[0, -3]], [1.0])], [[0, -1]], [1.0], [0], [2, 3], [14, 14])
    with pytest.raises(uhs.UhsError, match="outside the grid"):
        sweep([], [[2, 0]], [1.0], [-1], [2, 2], [15, 14])
    with pytest.raises(uhs.UhsError, match="source index"):
        sweep([([[0, 1]], [1.0])], [[0, 0]], [1.0], [1], [2, 2], [14, 14])
    with pytest.raises(uhs.UhsError, match="domain"):
        sweep([], [[0, 0]], [1.0], [-1], [2, 2], [2, 14])
    with pytest.raises(uhs.UhsError, match="distinct"):
        sweep([], [[0, 0]], [1.0], [-1], [2, 2], [14, 14], b=A)
    with pytest.raises(uhs.UhsError, match="producers"):
        sweep([([[0, 0]], [1.0])] * 5, [[0, 0]], [1.0], [0], [2, 2], [14, 14])
    with pytest.raises(uhs.UhsError, match="negative"):
        sweep([], [[0, 0]], [1.0], [-1], [2, 2], [14, 14], T=-1)
    with pytest.raises(uhs.UhsError, match="1 to 32"):
        sweep([], [[0, 0]] * 33, [1.0] * 33, [-1] * 33, [2, 2], [14, 14])
    with pytest.raises(uhs.UhsError, match="later producer"):
        sweep([([[0, 1]], [1.0], [0])], [[0, 0]], [1.0], [0], [2, 2], [14, 14])
    with pytest.raises(uhs.UhsError, match="outside the grid"):
        # the chain's reach (1 + 1 + 1 columns) exceeds the 2 left of the domain
        sweep([([[0, -1]], [1.0], [-1]), ([[0, -1]], [1.0], [0])], [[0, -1]], [1.0], [1], [2, 2], [14, 14])


def test_program_tuner_argument_errors(uhs):
    import ctypes
    A, B = HostBuf((20, 20)), HostBuf((20, 20))
    prog = uhs.ProgramSpec(2, [], [[0, 1]], [1.0], [-1], [1, 1], [19, 19])
    tile = (ctypes.c_int * 3)()
    L = uhs.lib()
    assert L.uhs_tune_program_f64(A.ptr, B.ptr, uhs.ext_array([20, 20]), ctypes.byref(prog.c), 0, tile,
                                  None, None, None) == -1
    assert L.uhs_tune_program_f64(A.ptr, B.ptr, uhs.ext_array([20, 20]), ctypes.byref(prog.c), 2, None,
                                  None, None, None) == -1


def test_program_kernel_choice_errors(uhs):
    """Tile and kernel checks happen before any launch."""
    A, B = HostBuf((90, 20, 70)), HostBuf((90, 20, 70))
    # 81 planes x 17 rows x 64 columns even for one-cell blocks: beyond 64 KiB.
    far = uhs.ProgramSpec(3, [], [[-40, 0, 0], [40, 0, 0], [0, -8, 0], [0, 8, 0]], [0.5, 0.5, 0.1, 0.1],
                          [-1, -1, -1, -1], [41, 9, 1], [49, 11, 69])
    near = uhs.ProgramSpec(3, [], [[-1, 0, 0], [1, 0, 0]], [0.5, 0.5], [-1, -1], [1, 1, 1], [89, 19, 69])
    shape = [90, 20, 70]
    with pytest.raises(uhs.UhsError, match="LDS") as e:
        uhs.program_sweep_f64(A, B, shape, far, 1, kernel=uhs.PROGRAM_TILED)
    assert e.value.code == -3
    with pytest.raises(uhs.UhsError, match="LDS"):
        uhs.program_sweep_f64(A, B, shape, near, 1, tile=[4096, 64, 4])
    with pytest.raises(uhs.UhsError, match="tile extents"):
        uhs.program_sweep_f64(A, B, shape, near, 1, tile=[0, 8, 4])
    with pytest.raises(uhs.UhsError, match="unknown program kernel"):
        uhs.program_sweep_f64(A, B, shape, near, 1, kernel=3)
    two_d = uhs.ProgramSpec(2, [], [[0, 1]], [1.0], [-1], [1, 1], [19, 69])
    with pytest.raises(uhs.UhsError, match="tile_z 1"):
        uhs.program_sweep_f64(A, B, [20, 70], two_d, 1, tile=[60, 8, 2])
    # Nothing launched for T = 0, whatever the choice.
    uhs.program_sweep_f64(A, B, shape, near, 0, kernel=uhs.PROGRAM_TILED, tile=[60, 8, 4])


ONE_SIDED_2D = """
stencil one_sided_small(double input[66][66]) {
    iteration(4)
    operation (k)
    kernel k {
        domain([2, 66][0, 64])
        expr {
            0.1*input[y-2][x] + 0.2*input[y-1][x] +
            0.3*input[y][x] + 0.4*input[y][x+1] + 0.5*input[y][x+2]
        }
    }
}
"""


def one_sided_check_loop(inp, T):
    """The reference driver's check loop for its one-sided example, restated
    in numpy: pairs of sweeps input -> tmp -> input over rows [2, M) and
    columns [0, N - 2), tmp starting at zero
    (UHStencil/examples/2d5pt_arbitrary_shape/stencil_2d5pt_arbitrary_shape_driver.serial.c:37-52)."""
    a = inp.copy()
    tmp = np.zeros_like(a)
    M, N = a.shape

    def sweep(src, dst):
        dst[2:M, 0:N - 2] = (0.1 * src[0:M - 2, 0:N - 2] + 0.2 * src[1:M - 1, 0:N - 2]
                             + 0.3 * src[2:M, 0:N - 2] + 0.4 * src[2:M, 1:N - 1] + 0.5 * src[2:M, 2:N])

    for _ in range(T // 2):
        sweep(a, tmp)
        sweep(tmp, a)
    return a


def test_one_sided_program_and_reference_check_loop(uhs, orc):
    prog, info = uhs.parse_dsl_program(ONE_SIDED_2D)
    assert info.nproducers == 0 and prog.lo == [2, 0] and prog.hi == [66, 64]
    assert prog.output.offsets.tolist() == [[-2, 0], [-1, 0], [0, 0], [0, 1], [0, 2]]
    rng = np.random.default_rng(9)
    A = rng.random((66, 66))
    B = np.zeros_like(A)
    want = one_sided_check_loop(A, 4)
    assert run_parsed(orc, prog, A, B, 4, PLAIN) == 0
    assert np.array_equal(A, want)
    # Without a domain, the written cells are those whose reads stay inside:
    # rows [2, 66), columns [0, 64).
    prog2, _ = uhs.parse_dsl_program(ONE_SIDED_2D.replace("domain([2, 66][0, 64])", ""))
    assert prog2.lo == [2, 0] and prog2.hi == [66, 64]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference example programs not present")
def test_reference_one_sided_examples(uhs, orc):
    prog, info = uhs.parse_dsl_program(path=os.path.join(REF, "2d5pt_arbitrary_shape", "stencil_2d5pt_arbitrary_shape.dsl"))
    assert info.iteration == 20 and list(info.grid[:2]) == [4098, 4098]
    assert prog.lo == [2, 0] and prog.hi == [4098, 4096] and info.nproducers == 0
    rng = np.random.default_rng(12)
    A = rng.random((4098, 4098))
    B = np.zeros_like(A)
    want = one_sided_check_loop(A, 2)
    assert run_parsed(orc, prog, A, B, 2, PLAIN) == 0
    assert np.array_equal(A, want)
    del A, B, want
    prog, info = uhs.parse_dsl_program(path=os.path.join(REF, "3d7pt_arbitrary_shape", "stencil_3d7pt_arbitrary_shape.dsl"))
    assert prog.lo == [0, 0, 0] and prog.hi == [256, 256, 256] and list(info.grid[:3]) == [258, 258, 258]
    assert prog.output.offsets.tolist() == [[2, 0, 0], [1, 0, 0], [0, 0, 0], [0, 0, 1], [0, 0, 2], [0, 1, 0], [0, 2, 0]]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference example programs not present")
def test_reference_nested_examples(uhs, orc):
    prog, info = uhs.parse_dsl_program(path=os.path.join(REF, "2d5pt_nested", "stencil_2d5pt_nested.dsl"))
    assert info.name.decode() == "stencil_2d5pt_nested" and info.operation.decode() == "stencil_2d9pt_x"
    assert info.nproducers == 1 and prog.sources.tolist() == [0] * 5
    assert prog.lo == [17, 17] and prog.hi == [4081, 4081] and info.iteration == 20
    # Two iterations of the driver's own check loop at the full 4098^2 grid.
    rng = np.random.default_rng(11)
    A = rng.random((4098, 4098))
    B = A.copy()
    want = in_place_check_loop(A, 2, 16)
    assert run_parsed(orc, prog, A, B, 2, PLAIN) == 0
    assert np.array_equal(A, want)
    del A, B, want

    prog, info = uhs.parse_dsl_program(path=os.path.join(REF, "3d7pt9pt_nested", "stencil_3d7pt9pt_nested.dsl"))
    assert info.operation.decode() == "stencil_3d9pt_x" and info.nproducers == 0
    assert prog.output.npoints == 9 and prog.lo == [5, 5, 5] and prog.hi == [253, 253, 253]
