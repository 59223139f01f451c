"""The autotuner's search logic on the CPU, through the mock timing backend
(uhs_tune_set_timer): timings and failures are injected by the test, the
grids are never touched, and no GPU is needed.  Checked: the winner is the
fastest candidate that ran (ties to the first), failed candidates are
recorded and skipped, the database makes a second search reuse every
timing, a launch failure stops the search and a rerun resumes after it, and
illegal candidates in a list never reach the timer.  The reference's loop is
a random search that records failing compiles and moves on
(scripts/evaluation_scripts_GPU/test_driver.py:198-256)."""
import ctypes

import pytest

E_CONFIG, E_HIP = -3, -4
FAKE_A, FAKE_B = ctypes.c_void_p(0x100000), ctypes.c_void_p(0x200000)
EXT = [96, 128, 160]


@pytest.fixture
def mock(uhs):
    yield uhs
    uhs.tune_set_timer(None)
    uhs.lib().uhs_tune_clear()


def fake_ms(c):
    """A deterministic timing per configuration, with ties."""
    return 1.0 + ((c.tile_x * 31 + c.tile_y * 17 + c.tile_z * 7 + c.vec * 3 + c.flags) % 97) / 10.0


def fails(c):
    return (c.tile_x + c.tile_y + c.tile_z) % 5 == 0


def search(u, st, db, cands=None):
    r = u.TuneResult()
    L = u.lib()
    dbp = str(db).encode() if db else None
    if cands is None:
        rc = L.uhs_tune_sweep(FAKE_A, FAKE_B, 3, u.ext_array(EXT), st.radius, ctypes.byref(st.c), u.DT_F64, 2,
                              dbp, ctypes.byref(r), None)
    else:
        arr = (u.Config * len(cands))(*cands)
        rc = L.uhs_tune_sweep_list(FAKE_A, FAKE_B, 3, u.ext_array(EXT), st.radius, ctypes.byref(st.c), u.DT_F64,
                                   arr, len(cands), 2, dbp, ctypes.byref(r), None)
    return rc, r


def expected_best(space, skip=()):
    best, best_ms = None, 1e30
    for c in space:
        if fails(c) or c.tuple() in skip:
            continue
        if fake_ms(c) < best_ms:
            best, best_ms = c, fake_ms(c)
    return best, best_ms


@pytest.mark.parametrize("case", ["3d1r_star", "3d2r_box"])
def test_search_skips_failures_and_installs_winner(mock, tmp_path, case):
    u = mock
    st = u.builtin(case)
    space = u.tune_space(EXT, st.radius, st)
    assert len(space) > 50
    calls = []

    def timer(c, reps):
        calls.append(c.tuple())
        assert reps == 2
        return (E_CONFIG, 0.0) if fails(c) else (0, fake_ms(c))

    u.tune_set_timer(timer)
    db = tmp_path / "tune.csv"
    rc, r = search(u, st, db)
    assert rc == 0
    nfail = sum(fails(c) for c in space)
    assert 0 < nfail < len(space)
    assert len(calls) == len(space) and len(set(calls)) == len(space)
    assert (r.evaluated, r.failed, r.reused) == (len(space) - nfail, nfail, 0)
    best, best_ms = expected_best(space)
    assert r.best.tuple() == best.tuple() and r.best_ms == pytest.approx(best_ms)
    # The winner is what cfg = NULL now uses for this problem.
    assert u.default_config(EXT, st.radius, st).tuple() == best.tuple()
    # One database line per candidate, failures with their status.
    lines = [l.split(",") for l in db.read_text().splitlines()]
    assert len(lines) == len(space)
    assert sum(int(l[-1]) == E_CONFIG for l in lines) == nfail


def test_search_resumes_from_database(mock, tmp_path):
    u = mock
    st = u.builtin("3d1r_star")
    space = u.tune_space(EXT, st.radius, st)
    db = tmp_path / "tune.csv"
    u.tune_set_timer(lambda c, reps: (E_CONFIG, 0.0) if fails(c) else (0, fake_ms(c)))
    rc, first = search(u, st, db)
    assert rc == 0

    def must_not_run(c, reps):
        raise AssertionError("timer called for %r" % (c,))

    u.tune_set_timer(must_not_run)
    rc, again = search(u, st, db)
    assert rc == 0
    assert again.reused == len(space) and again.best.tuple() == first.best.tuple()
    assert (again.evaluated, again.failed) == (first.evaluated, first.failed)


def test_launch_failure_stops_and_rerun_resumes(mock, tmp_path):
    u = mock
    st = u.builtin("3d1r_star")
    space = u.tune_space(EXT, st.radius, st)
    db = tmp_path / "tune.csv"
    k = len(space) // 3
    calls = []

    def timer(c, reps):
        calls.append(c.tuple())
        if len(calls) == k + 1:
            return (E_HIP, 0.0)
        return (E_CONFIG, 0.0) if fails(c) else (0, fake_ms(c))

    u.tune_set_timer(timer)
    rc, _ = search(u, st, db)
    assert rc == E_HIP and len(calls) == k + 1
    assert "stopped" in u.lib().uhs_last_error().decode()
    lines = db.read_text().splitlines()
    assert len(lines) == k + 1 and int(lines[-1].split(",")[-1]) == E_HIP
    crashed = calls[-1]

    calls.clear()
    u.tune_set_timer(lambda c, reps: (calls.append(c.tuple()), (E_CONFIG, 0.0) if fails(c) else (0, fake_ms(c)))[1])
    rc, r = search(u, st, db)
    assert rc == 0
    assert len(calls) == len(space) - (k + 1) and crashed not in calls
    assert r.reused == k + 1
    best, _ = expected_best(space, skip={crashed})
    assert r.best.tuple() == best.tuple()


def test_illegal_candidates_never_reach_the_timer(mock, tmp_path):
    u = mock
    st = u.builtin("3d1r_star")
    legal = u.tune_space(EXT, st.radius, st)[:5]
    illegal = [u.Config(0, 8, 64, 1, 0), u.Config(128, 16, 64, 3, 0), u.Config(64, 8, 64, 2, 1 << 20)]
    calls = []
    u.tune_set_timer(lambda c, reps: (calls.append(c.tuple()), (0, fake_ms(c)))[1])
    rc, r = search(u, st, None, cands=illegal + list(legal))
    assert rc == 0
    assert sorted(calls) == sorted(c.tuple() for c in legal)
    assert (r.evaluated, r.failed) == (5, 3)
    # Only illegal candidates: the search reports that nothing ran.
    calls.clear()
    rc, _ = search(u, st, None, cands=illegal)
    assert rc == E_CONFIG and not calls
