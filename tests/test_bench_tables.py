"""CPU: the bench line's counters come from the committed profile tables,
averaged over exactly the updates a run times (--warmup 5 --steps 20 times
updates 6..25), and the tables name the workload and shape they were
measured on."""
import os
import re

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EVALS = [2_000_000_000] * 20


def _rows(path):
    out = {}
    for line in open(path):
        m = re.match(r"\|\s*(\d+)\s*\|(.*)\|\s*$", line)
        if m:
            out[int(m.group(1))] = [float(c) for c in m.group(2).split("|")]
    return out


def test_traffic_and_issue_average_the_timed_updates():
    rows = _rows(bench.PROFILE_TABLE)
    c = bench.profile_counters(bench.PROFILE_TABLE, 5, 20, EVALS, "map0-ppw16-spatial-tile256", seed=7, n=100000,
                               m=20000)
    upd = range(6, 26)
    assert abs(c["traffic"] - sum(rows[u][0] + rows[u][1] for u in upd) / 20) < 1e-6 * c["traffic"]
    assert abs(c["valu_per_64_evals"] - sum(rows[u][2] for u in upd) / (sum(EVALS) / 64)) < 1e-9
    assert "updates 6..25" in c["note"]


def test_tables_refuse_another_workload_or_range():
    c = bench.profile_counters(bench.PROFILE_TABLE, 5, 20, EVALS, "map0-ppw32-spatial-tile256", seed=7, n=100000,
                               m=20000)
    assert "traffic" not in c and "is for" in c["note"]
    c = bench.profile_counters(bench.PROFILE_TABLE, 100, 20, EVALS, "map0-ppw16-spatial-tile256", seed=7, n=100000,
                               m=20000)
    assert "traffic" not in c and "covers updates" in c["note"]
    assert bench.profile_clock(bench.CLOCK_TABLE, 100, 20) is None


def test_clock_and_issue_at_that_clock():
    rows = _rows(bench.CLOCK_TABLE)
    clk = bench.profile_clock(bench.CLOCK_TABLE, 5, 20)
    assert abs(clk - sum(rows[u][2] for u in range(6, 26)) / 20) < 1e-9
    assert 1.5 < clk <= 2.4
    issue = bench.issue_block({"valu_per_64_evals": 36.0}, 2e9, 2.5, clk)
    assert abs(issue["frac"] * 2.4 - issue["frac_at_profiled_clock"] * clk) < 1e-9
    assert bench.profiled_shape(bench.PROFILE_TABLE) == "0,16,1,256"
