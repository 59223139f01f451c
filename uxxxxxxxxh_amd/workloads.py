"""Inputs of the benchmark workloads, shared by bench.py and the tests so both
build exactly the same grids.

* uneven_batch(): 64 independent 3D FP64 grids, every axis extent drawn from
  {64, 96, ..., 384} (seeded), grids 0-31 the 7-point star (3d1r_star) and
  32-63 the 125-point box (3d2r_box), each with its own seeded symmetric
  coefficients normalised to sum 1, interiors U[0,1), zero halos.
* dyadic_coeffs(): symmetric coefficients that are powers of two (the centre
  term takes the remainder) and sum to exactly 1, so every product and
  partial sum of a sweep over a linear field of integers is exact: the field
  is then a fixed point of the sweep, bit for bit, for any number of steps.
* linear_field(): 1 + x + 2y (+ 4z), halo included.
"""
import numpy as np

BATCH_EXTENTS = tuple(range(64, 385, 32))


def uneven_batch(n=64, seed=2024):
    """[{case, coeff_seed, ext, seed_a, seed_b}] for the benchmark batch."""
    rng = np.random.default_rng(seed)
    choices = list(BATCH_EXTENTS)
    out = []
    for i in range(n):
        case = "3d1r_star" if i < n // 2 else "3d2r_box"
        ext = [int(rng.choice(choices)) for _ in range(3)]
        out.append({"case": case, "coeff_seed": 500 + i, "ext": ext, "seed_a": 7 + i, "seed_b": 1007 + i})
    return out


def dyadic_coeffs(offsets):
    """2^-m for every off-centre term and 1 - (P-1) 2^-m for the centre, with
    2^m >= P: positive, symmetric, summing to exactly 1."""
    off = np.asarray(offsets).reshape(len(offsets), -1)
    P = len(off)
    m = int(np.ceil(np.log2(P)))
    c = np.full(P, 2.0 ** -m)
    centre = [t for t in range(P) if not np.any(off[t])]
    assert len(centre) == 1, "stencil without a single centre term"
    c[centre[0]] = 1.0 - (P - 1) * 2.0 ** -m
    assert c.sum() == 1.0 and np.all(c > 0)
    return c


def linear_field(shape, dtype=np.float64):
    """1 + x + 2y + 4z over the halo-inclusive grid (x the unit-stride axis)."""
    idx = np.indices(shape, dtype=np.float64)
    w = [4.0, 2.0, 1.0][-len(shape):]
    f = np.ones(shape)
    for k in range(len(shape)):
        f += w[k] * idx[k]
    return f.astype(dtype)
