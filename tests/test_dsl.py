"""The DSL reader (uhs_parse_dsl) against the stencil tables of the 12
benchmark DSLs and on hand-written programs, including the error paths."""
import glob
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = json.load(open(os.path.join(HERE, "golden", "shapes.json")))
REF = "/root/reference/UH-GPU"


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference DSL files not present")
@pytest.mark.parametrize("case", list(SHAPES))
def test_reference_dsl_files(uhs, case):
    path = glob.glob(os.path.join(REF, case, "*.dsl"))[0]
    st, info = uhs.parse_dsl(path=path)
    s = SHAPES[case]
    assert info.name.decode() == s["symbol"]
    assert st.offsets.tolist() == s["offsets"] and st.coeffs.tolist() == s["coeffs"]
    assert info.iteration == s["iteration"] and info.halo == s["radius"]
    assert list(info.grid[:s["dims"]]) == s["grid"]
    assert [[info.domain_lo[k], info.domain_hi[k]] for k in range(s["dims"])] == s["domain"]
    assert uhs.plan_kind(list(info.extents[:s["dims"]]), info.halo, st) == (
        "win2d" if s["dims"] == 2 else ("star3d" if s["kind"] == "star" else "box3d"))


PROGRAM = """
# a one-sided 3D stencil with mixed term forms
stencil stencil_upwind(double field[20][22][24]) {
    iteration(2)
    operation (k)
    kernel k {
        tile(4, 4, 8)
        domain([2,18][2,20][2,22])
        expr {
            -0.25*field[z][y][x] + field[z-1][y][x]*0.5
            - 0.125*field[z][y-2][x+1] + field[z+2][y][x-2]
        }
    }
}
"""


def test_hand_written_program(uhs):
    st, info = uhs.parse_dsl(PROGRAM)
    assert info.name.decode() == "stencil_upwind" and info.kernel.decode() == "k"
    assert info.dims == 3 and info.iteration == 2 and info.halo == 2 and not info.is_float
    assert list(info.extents) == [16, 18, 20]
    assert st.offsets.tolist() == [[0, 0, 0], [-1, 0, 0], [0, -2, 1], [2, 0, -2]]
    assert st.coeffs.tolist() == [-0.25, 0.5, -0.125, 1.0]
    assert uhs.plan_kind([16, 18, 20], 2, st) == "program"


def test_float_program_and_defaults(uhs):
    text = "stencil s(float a[10][12]) { kernel k { expr { 0.5*a[y][x-1] + 0.5*a[y][x+1] } } }"
    st, info = uhs.parse_dsl(text)
    assert info.is_float == 1 and info.iteration == 1 and info.halo == 1
    assert list(info.extents[:2]) == [8, 10]


@pytest.mark.parametrize("text,code", [
    ("stencil s(double a[10][10]) { kernel k { expr { (0.5*a[y][x]) } } }", -7),
    ("stencil s(double a[10][10]) { kernel k { expr { a[y][x]/2.0 } } }", -7),
    ("stencil s(double a[10][10]) { iteration(3) kernel k { expr { a[y][x] } } }", -7),
    ("stencil s(double a[10][10]) { kernel k { expr { a[y][x][x] } } }", -7),
    ("stencil s(double a[10][10]) { kernel k { expr { b[y][x] } } }", -7),
    ("stencil s(double a[10][10]) { kernel k { expr { a[y][x] } } kernel k2 { expr { a[y][x] } } }", -2),
    ("stencil s(double a[10][10]) { kernel k { domain([1,9][2,8]) expr { a[y][x+1] } } }", -2),
    ("stencil s(double a[10][10]) { kernel k { expr { a[y][x] }", -7),
])
def test_rejections(uhs, text, code):
    with pytest.raises(uhs.UhsError) as e:
        uhs.parse_dsl(text)
    assert e.value.code == code


def test_missing_file(uhs):
    with pytest.raises(uhs.UhsError) as e:
        uhs.parse_dsl(path="/nonexistent/x.dsl")
    assert e.value.code == -6
