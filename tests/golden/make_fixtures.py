"""Write the golden stencil tables and small known-answer vectors.

Run in a container where the reference tree is mounted (it reads the 12
benchmark DSL files under UH-GPU/<case>/*.dsl); the outputs are committed so
nothing at test time needs the reference.

* ``shapes.json``: for each UH-GPU case, the stencil name, dims, radius,
  star/box class, term offsets and coefficients in DSL term order, the
  domain, the grid extents (halo included) and iteration(T), read from the
  DSL text (UH-GPU/<case>/*.dsl; positional offsets as in
  UHStencil/include/Parser/Parser.h:525-575).
* ``kat_<case>.npz``: tiny halo-inclusive grids A, B (every cell random,
  halo included, so halo handling is pinned) and the expected A, B after
  T = 1, 2 and 3 ping-pong sweeps (MLIRGen.cpp:317-349,
  ConvertStencilToStandard.cpp:298-337).  Expected values come from
  evaluating the DSL ``expr`` text itself, point by point, with Python's
  own parser: '+' and '*' have the DSL's precedence and left associativity
  (Parser.h:437-468,591-606) and Python floats are IEEE doubles without
  contraction, so this is the reference's sum in its own term order.  Each
  file holds the vectors twice: once with the DSL's coefficients and once
  with symmetric normalised random coefficients substituted into the same
  text.
"""
import json
import os
import re
import sys

import numpy as np

REF = "/root/reference/UH-GPU"
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["2d1r_star", "2d2r_star", "2d3r_star", "2d1r_box", "2d2r_box", "2d3r_box",
         "3d1r_star", "3d2r_star", "3d3r_star", "3d1r_box", "3d2r_box", "3d3r_box"]
TERM = re.compile(r"([0-9]*\.?[0-9]+)\s*\*\s*input((?:\[[^\]]*\])+)")
BRACKET = re.compile(r"\[\s*[a-z]\s*(?:([+-])\s*([0-9]+))?\s*\]")


def parse_dsl(text):
    name = re.search(r"stencil\s+(\w+)\s*\(\s*double\s+input((?:\[\d+\])+)", text)
    grid = [int(g) for g in re.findall(r"\[(\d+)\]", name.group(2))]
    iteration = int(re.search(r"iteration\s*\(\s*(\d+)\s*\)", text).group(1))
    domain = [(int(a), int(b)) for a, b in
              re.findall(r"\[\s*(\d+)\s*,\s*(\d+)\s*\]", re.search(r"domain\s*\(([^)]*)\)", text).group(1))]
    expr = re.search(r"expr\s*\{(.*?)\}", text, re.S).group(1)
    offs, coefs = [], []
    for c, br in TERM.findall(expr):
        o = []
        for sign, mag in BRACKET.findall(br):
            o.append(0 if not mag else (int(mag) if sign == "+" else -int(mag)))
        offs.append(o)
        coefs.append(float(c))
    return dict(symbol=name.group(1), grid=grid, iteration=iteration, domain=domain,
                expr=expr, offsets=offs, coeffs=coefs)


def mix64(z):
    m = (1 << 64) - 1
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def u01(seed, i):
    return (mix64((seed + (i + 1) * 0x9E3779B97F4A7C15) & ((1 << 64) - 1)) >> 11) * 2.0 ** -53


def sym_coeffs(offs, seed):
    """Symmetric U[0,1) coefficients normalised to 1, draws in term order."""
    c, draws = [], 0
    for t, o in enumerate(offs):
        partner = next((s for s in range(t) if all(a == -b for a, b in zip(offs[s], o))), None)
        if partner is None:
            c.append(u01(seed, draws)); draws += 1
        else:
            c.append(c[partner])
    s = 0.0
    for v in c:
        s += v
    return [v / s for v in c]


def compile_expr(expr, coefs=None):
    """Turn the DSL expr text into a Python expression over G(o0,o1[,o2]).

    With `coefs`, the t-th literal coefficient is replaced by K[t]."""
    counter = [0]

    def term(m):
        t = counter[0]; counter[0] += 1
        o = []
        for sign, mag in BRACKET.findall(m.group(2)):
            o.append("0" if not mag else (mag if sign == "+" else "-" + mag))
        c = m.group(1) if coefs is None else "K[%d]" % t
        return "%s*G(%s)" % (c, ",".join(o))
    src = TERM.sub(term, expr)
    return compile(" ".join(src.split()), "<dsl>", "eval")


def sweep_by_text(code, K, src, dst, r):
    dims = src.ndim
    out = dst.copy()
    inner = [range(r, n - r) for n in src.shape]
    if dims == 2:
        for i in inner[0]:
            for j in inner[1]:
                out[i, j] = eval(code, {"G": lambda a, b: float(src[i + a, j + b]), "K": K})
    else:
        for i in inner[0]:
            for j in inner[1]:
                for k in inner[2]:
                    out[i, j, k] = eval(code, {"G": lambda a, b, c: float(src[i + a, j + b, k + c]), "K": K})
    return out


def main():
    shapes = {}
    for case in CASES:
        d = os.path.join(REF, case)
        dsl = [f for f in os.listdir(d) if f.endswith(".dsl")][0]
        with open(os.path.join(d, dsl)) as fh:
            info = parse_dsl(fh.read())
        dims = len(info["grid"])
        r = max(abs(v) for o in info["offsets"] for v in o)
        star = all(sum(1 for v in o if v) <= 1 for o in info["offsets"])
        shapes[case] = dict(case=case, dsl_file="UH-GPU/%s/%s" % (case, dsl), symbol=info["symbol"],
                            dims=dims, radius=r, kind="star" if star else "box",
                            npoints=len(info["offsets"]), offsets=info["offsets"],
                            coeffs=info["coeffs"], iteration=info["iteration"],
                            grid=info["grid"], domain=info["domain"])
        rng = np.random.default_rng(1234 + CASES.index(case))
        n = [10, 12] if dims == 2 else [5, 6, 7]
        full = [v + 2 * r for v in n]
        A0 = rng.random(full)
        B0 = rng.random(full)
        vec = {"A0": A0, "B0": B0, "interior": np.array(n), "halo": np.array(r)}
        for tag, K in (("dsl", None), ("sym", sym_coeffs(info["offsets"], 77))):
            code = compile_expr(info["expr"], K)
            A, B = A0.copy(), B0.copy()
            for T in (1, 2, 3):
                if T % 2:
                    B = sweep_by_text(code, K, A, B, r)
                else:
                    A = sweep_by_text(code, K, B, A, r)
                vec["%s_A_T%d" % (tag, T)] = A.copy()
                vec["%s_B_T%d" % (tag, T)] = B.copy()
            if K is not None:
                vec["sym_coeffs"] = np.array(K)
        np.savez_compressed(os.path.join(HERE, "kat_%s.npz" % case), **vec)
        print(case, shapes[case]["symbol"], shapes[case]["npoints"], "terms", file=sys.stderr)
    with open(os.path.join(HERE, "shapes.json"), "w") as fh:
        json.dump(shapes, fh, indent=1)


if __name__ == "__main__":
    main()
