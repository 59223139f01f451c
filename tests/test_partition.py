"""Splitting a batch over several GPUs (uhs_batch_partition): longest
processing time first.  Host only.

Pinned by: brute force over every assignment on small instances (the
split is within 4/3 - 1/(3m) of the best, Graham's bound for LPT); the
classic instance where LPT is not optimal (m = 2, costs 3,3,2,2,2: LPT
gives 7, the best is 6); the tie rule; and the cost estimate against
points x max(1, terms / 64)."""
import itertools

import numpy as np
import pytest


def loads_of(gpu, cost, m):
    out = [0.0] * m
    for g, c in zip(gpu, cost):
        out[g] += c
    return out


def grids_for(u, n, seed=0):
    rng = np.random.default_rng(seed)
    kinds = ["3d1r_star", "3d2r_box", "3d1r_box", "3d3r_star"]
    out = []
    for i in range(n):
        st = u.builtin(kinds[i % len(kinds)])
        out.append(([int(v) for v in rng.integers(16, 200, 3)], st.radius, st))
    return out


def test_every_grid_once_and_loads_sum(uhs):
    grids = grids_for(uhs, 23)
    cost = list(np.random.default_rng(1).uniform(0.1, 10.0, 23))
    for m in (1, 2, 3, 8):
        gpu, load = uhs.batch_partition(grids, m, cost)
        assert len(gpu) == 23 and all(0 <= g < m for g in gpu)
        assert np.allclose(load, loads_of(gpu, cost, m))


@pytest.mark.parametrize("seed", range(20))
def test_within_graham_bound_of_brute_force(uhs, seed):
    rng = np.random.default_rng(100 + seed)
    n, m = int(rng.integers(3, 9)), int(rng.integers(2, 4))
    cost = [float(v) for v in rng.integers(1, 20, n)]
    grids = grids_for(uhs, n)
    gpu, load = uhs.batch_partition(grids, m, cost)
    best = min(max(loads_of(a, cost, m)) for a in itertools.product(range(m), repeat=n))
    assert max(load) <= (4.0 / 3.0 - 1.0 / (3 * m)) * best + 1e-9
    assert max(load) >= best


def test_classic_lpt_instance(uhs):
    # LPT puts 3 | 3, then 2 | 2, then the last 2 on a tie -> loads 7 and 5;
    # the best split is 3+3 | 2+2+2 = 6.
    gpu, load = uhs.batch_partition(grids_for(uhs, 5), 2, [3.0, 3.0, 2.0, 2.0, 2.0])
    assert sorted(load) == [5.0, 7.0]
    assert gpu == [0, 1, 0, 1, 0]


def test_ties_go_to_the_lower_gpu_in_grid_order(uhs):
    gpu, load = uhs.batch_partition(grids_for(uhs, 6), 3, [1.0] * 6)
    assert gpu == [0, 1, 2, 0, 1, 2] and load == [2.0, 2.0, 2.0]


def test_estimated_costs(uhs):
    grids = grids_for(uhs, 9, seed=3)
    weight = [float(np.prod(ext)) * max(1.0, st.npoints / 64.0) for ext, _, st in grids]
    assert uhs.batch_partition(grids, 4) == uhs.batch_partition(grids, 4, weight)


def test_argument_errors(uhs):
    grids = grids_for(uhs, 3)
    for bad in (lambda: uhs.batch_partition(grids, 0),
                lambda: uhs.batch_partition(grids, 2, [1.0, -1.0, 2.0]),
                lambda: uhs.batch_partition(grids, 2, [1.0, float("nan"), 2.0])):
        with pytest.raises(uhs.UhsError) as e:
            bad()
        assert e.value.code == -1
    odd = uhs.StencilSpec(3, [[0, 0, 0], [1, 0, 0]], [0.5, 0.5])
    with pytest.raises(uhs.UhsError) as e:
        uhs.batch_partition([([8, 8, 8], 1, odd)], 2)
    assert e.value.code == -2
    assert uhs.batch_partition([([8, 8, 8], 1, odd)], 2, [1.0]) == ([0], [1.0, 0.0])
    # a grid without offsets cannot be estimated (rejected, not dereferenced)
    arr = (uhs.GridDesc * 1)()
    arr[0] = uhs.GridDesc(None, None, 3, (uhs.ctypes.c_int64 * 3)(8, 8, 8), 1, uhs.Stencil(3, 7, None, None))
    gpu = (uhs.ctypes.c_int * 1)()
    assert uhs.lib().uhs_batch_partition(arr, 1, 2, None, gpu, None) == -2
