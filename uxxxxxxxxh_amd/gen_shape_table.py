"""Emit csrc/shape_table.inc: the 12 benchmark stencils as constexpr tables.

Input is the stencil table extracted from the reference's DSL files
(UH-GPU/<case>/*.dsl, one `expr` block each; offsets are positional,
UHStencil/include/Parser/Parser.h:525-575) -- tests/golden/shapes.json.
Each stencil keeps its DSL term order, which fixes the summation order of
the sweep (Parser.h:437-468).

    python uxxxxxxxxh_amd/gen_shape_table.py tests/golden/shapes.json
"""
import json
import os
import sys

ORDER = ["2d1r_star", "2d2r_star", "2d3r_star", "2d1r_box", "2d2r_box", "2d3r_box",
         "3d1r_star", "3d2r_star", "3d3r_star", "3d1r_box", "3d2r_box", "3d3r_box"]


def main(path):
    with open(path) as fh:
        shapes = json.load(fh)
    out = ["// Generated by uxxxxxxxxh_amd/gen_shape_table.py -- do not edit.",
           "// The 12 UH-GPU benchmark stencils in DSL term order (UH-GPU/<case>/*.dsl).",
           "#pragma once", ""]
    for i, case in enumerate(ORDER):
        s = shapes[case]
        d, P = s["dims"], s["npoints"]
        offs = ",".join("{%s}" % ",".join(str(v) for v in o) for o in s["offsets"])
        coefs = ",".join(repr(float(c)) for c in s["coeffs"])
        flat = ",".join(str(v) for o in s["offsets"] for v in o)
        out += ["struct Prog_%s {" % case,
                "  static constexpr int ID = %d, DIMS = %d, R = %d, P = %d;" % (i, d, s["radius"], P),
                "  static constexpr bool STAR = %s;" % ("true" if s["kind"] == "star" else "false"),
                "  static constexpr int OFF[%d][%d] = {%s};" % (P, d, offs),
                "};",
                "static const int kOff_%s[%d] = {%s};" % (case, P * d, flat),
                "static const double kCoef_%s[%d] = {%s};" % (case, P, coefs), ""]
    out.append("#define UHS_FOR_EACH_BUILTIN(X) \\")
    out += ["  X(%s) \\" % c for c in ORDER[:-1]] + ["  X(%s)" % ORDER[-1], ""]
    out.append("static const BuiltinStencil kBuiltins[%d] = {" % len(ORDER))
    for i, case in enumerate(ORDER):
        s = shapes[case]
        out.append('  {"%s", "%s", %d, %d, %d, %s, kOff_%s, kCoef_%s},' % (
            case, s["symbol"], s["dims"], s["radius"], s["npoints"],
            "1" if s["kind"] == "star" else "0", case, case))
    out.append("};")
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "shape_table.inc")
    with open(dst, "w") as fh:
        fh.write("\n".join(out) + "\n")
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "tests/golden/shapes.json")
