"""rocprofv3 summary helper (host only): kernel names as the profiles and
bench.py key them, and per-dispatch counter sums from a results database."""
import sqlite3

from loquat_amd import profsummary as ps


def test_short_name_keeps_template_arguments_only():
    k = ("void lq::(anonymous namespace)::k_interaction_v3<lq::DescTable, 2, true, false, true, 40, 4, 1, 16, 2>"
         "(lq::DescTable, lq::DevState, lq::DevRates, int const*, int const*, lq::IndexMap, int, int*, lq::StaleIn)")
    assert ps.short_name(k) == "k_interaction_v3<DescTable, 2, true, false, true, 40, 4, 1, 16, 2>"
    assert ps.short_name("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def test_counters_sum_the_rows_of_one_dispatch(tmp_path):
    d = tmp_path / "run"
    d.mkdir()
    con = sqlite3.connect(d / "run_results.db")
    con.execute("create table counters_collection (dispatch_id int, kernel_name text, counter_name text, value real)")
    rows = [(1, "void k_a<1>(int)", "FETCH_SIZE", 10.0), (1, "void k_a<1>(int)", "FETCH_SIZE", 5.0),
            (2, "void k_a<1>(int)", "FETCH_SIZE", 7.0), (3, "__amd_rocclr_copyBuffer", "FETCH_SIZE", 100.0)]
    con.executemany("insert into counters_collection values (?, ?, ?, ?)", rows)
    con.commit()
    c = ps.counters(str(tmp_path))
    assert sorted(c["k_a<1>"]["FETCH_SIZE"]) == [7.0, 15.0]
    assert c["__amd_rocclr_copyBuffer"]["FETCH_SIZE"] == [100.0]
