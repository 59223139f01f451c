"""Host sanitizer build of the DSL reader: tests/asan/dsl_fuzz.cpp and
uxxxxxxxxh_amd/csrc/dsl.cpp compiled with g++ -fsanitize=address,undefined
(no GPU code in the build), run on the 12 benchmark stencils written out as
DSL text from the stencil table, on the nested and one-sided programs, and on
every prefix and 300 seeded one-byte mutations of each.  Any out-of-bounds
access, overflow, undefined operation or leak in the reader aborts the run;
the whole texts must still parse."""
import json
import os
import shutil
import subprocess

import pytest

from tests.test_program import NESTED_2D, ONE_SIDED_2D

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SHAPES = json.load(open(os.path.join(HERE, "golden", "shapes.json")))
AXES = {2: "yx", 3: "zyx"}


def dsl_text(s):
    """A benchmark stencil as DSL text: declaration, iteration, tile, domain
    and the expr in term order, as the reference's .dsl files write them."""
    dims = s["dims"]
    ax = AXES[dims]
    decl = "".join("[%d]" % g for g in s["grid"])
    dom = "".join("[%d,%d]" % tuple(d) for d in s["domain"])
    terms = []
    for off, c in zip(s["offsets"], s["coeffs"]):
        idx = ""
        for a, o in zip(ax, off):
            idx += "[%s%s]" % (a, "" if o == 0 else ("%+d" % o))
        terms.append("%r*B%s" % (c, idx))
    expr = "\n            + ".join(terms)
    tile = ", ".join(["4"] * dims)
    return ("# %s\nstencil %s(double B%s) {\n    iteration(%d)\n    operation(k)\n    kernel k {\n"
            "        tile(%s)\n        domain(%s)\n        expr {\n            %s\n        }\n    }\n}\n"
            % (s["case"], s["symbol"], decl, s["iteration"], tile, dom, expr))


@pytest.fixture(scope="module")
def fuzz_bin(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not present")
    out = tmp_path_factory.mktemp("asan") / "dsl_fuzz"
    cmd = [gxx, "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "uxxxxxxxxh_amd", "csrc"),
           os.path.join(HERE, "asan", "dsl_fuzz.cpp"),
           os.path.join(ROOT, "uxxxxxxxxh_amd", "csrc", "dsl.cpp"), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "hip/hip_runtime.h" in r.stderr:
        pytest.skip("ROCm headers not present")
    assert r.returncode == 0, r.stderr[-3000:]
    return str(out)


def run_fuzz(fuzz_bin, tmp_path, texts, mutations=300):
    files = []
    for i, t in enumerate(texts):
        p = tmp_path / ("in%02d.dsl" % i)
        p.write_text(t)
        files.append(str(p))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([fuzz_bin, str(mutations)] + files, capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    rows = [list(map(int, line.split())) for line in r.stdout.split("\n") if line.strip()]
    assert len(rows) == len(texts)
    return rows


def test_benchmark_stencils_under_sanitizers(fuzz_bin, tmp_path):
    texts = [dsl_text(SHAPES[c]) for c in SHAPES]
    rows = run_fuzz(fuzz_bin, tmp_path, texts)
    for (whole, prefixes, muts, _), c, t in zip(rows, SHAPES, texts):
        # The stencil reader takes every one; the program reader those of at
        # most 32 terms (its kernels' limit) and refuses the others.
        assert whole == (3 if SHAPES[c]["npoints"] <= 32 else 1), t
        assert prefixes == len(t) and muts == 300


def test_programs_under_sanitizers(fuzz_bin, tmp_path):
    texts = [NESTED_2D, ONE_SIDED_2D]
    rows = run_fuzz(fuzz_bin, tmp_path, texts)
    # The nested program has two kernels: only the program reader takes it.
    assert rows[0][0] == 2
    assert rows[1][0] & 2
