"""The kernel-trace gap summary on a synthetic trace with known idle time."""
import sqlite3

import pytest

from loquat_amd.tracegaps import step_gaps


def _trace():
    # three steps of 100 us: begin (1 us), a (50 us), gap 4 us, b (40 us),
    # idle 5 us before the next step; a fourth k_step_begin closes step 3
    rows, t = [], 0
    for _ in range(3):
        rows += [("k_step_begin", t, t + 1000), ("k_a", t + 1000, t + 51000), ("k_b", t + 55000, t + 95000)]
        t += 100000
    rows.append(("k_step_begin", t, t + 1000))
    return rows


def test_step_gaps_known_idle():
    r = step_gaps(_trace()[::-1])  # any order
    assert r["steps"] == 3 and r["kernels_per_step"] == 3
    assert r["span_us"] == pytest.approx(100.0) and r["busy_us"] == pytest.approx(91.0)
    assert r["idle_us"] == pytest.approx(9.0) and r["largest_gap_us"] == pytest.approx(5.0)
    assert step_gaps(_trace(), skip=2)["steps"] == 1
    assert step_gaps(_trace(), skip=1, nsteps=1)["steps"] == 1
    with pytest.raises(ValueError):
        step_gaps(_trace(), skip=3)


def test_reads_the_kernels_view(tmp_path):
    db = tmp_path / "run_results.db"
    con = sqlite3.connect(db)
    con.execute("create table k (name text, start integer, end integer)")
    con.executemany("insert into k values (?, ?, ?)", _trace())
    con.execute("create view kernels as select name, start, end from k")
    con.commit()
    rows = list(sqlite3.connect(db).execute("select name, start, end from kernels"))
    assert step_gaps(rows)["idle_us"] == pytest.approx(9.0)
