"""The reference-workload driver (uxxxxxxxxh_amd/drivers/uhs_ref_driver) for
all 12 UH-GPU cases at the reference's own sizes (8192^2, 512^3): glibc
rand() % 50 over every cell, the case's reference symbol from its shim
library, iteration(4), the timing print, and the driver's own serial check
(|cpu - gpu| <= 1e-6, UH-GPU/3d1r_star/3d7pt_main.cpp:84-96).  The test then
recomputes the driver's input (srand(1) restarts glibc's sequence where a
fresh process starts it) and the oracle's fused sweep, and compares the two
checksums of the result grid the driver prints: bitwise equality."""
import ctypes
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from tests.oracle_py import FUSED

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SHAPES = json.load(open(os.path.join(HERE, "golden", "shapes.json")))
DRIVER = os.path.join(ROOT, "uxxxxxxxxh_amd", "drivers", "uhs_ref_driver")


def checksums(a):
    """Sum and position-weighted sum of the bit patterns, mod 2^64."""
    v = a.reshape(-1).view(np.uint64)
    s = w = np.uint64(0)
    chunk = 1 << 24
    with np.errstate(over="ignore"):
        for i in range(0, v.size, chunk):
            blk = v[i:i + chunk]
            idx = np.arange(i + 1, i + 1 + blk.size, dtype=np.uint64)
            s = s + blk.sum(dtype=np.uint64)
            w = w + (blk * idx).sum(dtype=np.uint64)
    return int(s), int(w)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("case", list(SHAPES))
def test_reference_driver_full_size(gpu, orc, case):
    assert os.path.exists(DRIVER), "driver not built (build())"
    print(case, "reference driver, full size", file=sys.stderr, flush=True)
    p = subprocess.run([DRIVER, case], capture_output=True, text=True, timeout=500)
    print(p.stdout, p.stderr, file=sys.stderr, flush=True)
    assert p.returncode == 0, p.stderr
    out = p.stdout
    assert "device elapsed Time:" in out and "cpu elapsed Time:" in out
    assert "Verify Success" in out
    m = re.search(r"result checksum: sum ([0-9a-f]{16}) weighted ([0-9a-f]{16})", out)
    assert m
    s = SHAPES[case]
    r = s["radius"]
    n = 8192 if s["dims"] == 2 else 512
    shape = [n + 2 * r] * s["dims"]
    A = np.empty(shape)
    ctypes.CDLL(None).srand(1)
    orc.L.orc_fill_rand50_f64(A, A.size)
    B = A.copy()
    orc.iterate(A, B, [n] * s["dims"], r, s["offsets"], s["coeffs"], s["iteration"], FUSED)
    assert checksums(A) == (int(m.group(1), 16), int(m.group(2), 16))


def test_driver_rejects_unknown_case():
    assert os.path.exists(DRIVER)
    p = subprocess.run([DRIVER, "4d1r_star"], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0 and "unknown case" in p.stderr
