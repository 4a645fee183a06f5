"""The rocprofv3 summary tool (simplenavierstokessolver_amd/profiling.py) on
small synthetic CSV files: kernel groups, the copy-calibrated traffic and the
counter table with its L2 hit rate."""
import csv

import pytest

from simplenavierstokessolver_amd import profiling as prof


def write_counters(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writerows(rows)


PRED = "void sns::k_pred_rows<512, 8, true, sns::ArgOne<sns::PredArgs> >(sns::ArgOne<sns::PredArgs>)"
COPY = "void sns::k_copy(double const*, double*, long)"


def test_groups():
    assert prof.group_of(PRED) == "pred"
    assert prof.group_of("void sns::k_pfwd_a_p2<64, sns::ArgOne<sns::PoisArgs> >(x)") == \
        "poisson_fwd"
    assert prof.group_of("void sns::k_yapply_rows<1, S>(S)") == "ysolve"
    assert prof.group_of("void sns::k_yreduced_c<16, 4, S>(S)") == "ysolve"


def test_traffic_calibration(tmp_path):
    """Counts are scaled by what the copy kernel reads and writes: a copy of
    1 MiB that reports half its bytes fetched gives a fetch factor of 2."""
    fe, wr = tmp_path / "f.csv", tmp_path / "w.csv"
    write_counters(fe, [(COPY, "FETCH_SIZE", 512.0), (PRED, "FETCH_SIZE", 100.0),
                        (PRED, "FETCH_SIZE", 300.0)])
    write_counters(wr, [(COPY, "WRITE_SIZE", 1024.0), (PRED, "WRITE_SIZE", 50.0)])
    t = prof.traffic([str(fe)], [str(wr)], 1 << 20, 10)
    assert t["fetch_factor"] == 2.0 and t["write_factor"] == 1.0
    (row,) = t["kernels"]
    assert row["kernel"] == "k_pred_rows" and row["dispatches"] == 2
    assert row["fetch_bytes"] == pytest.approx(200.0 * 1024 * 2)
    assert row["write_bytes"] == pytest.approx(50.0 * 1024)


def test_counter_table_and_hit_rate(tmp_path, capsys):
    p = tmp_path / "c.csv"
    write_counters(p, [(PRED, "TCC_HIT_sum", 70.0), (PRED, "TCC_MISS_sum", 30.0),
                       (PRED, "TCC_HIT_sum", 90.0), (PRED, "TCC_MISS_sum", 10.0)])
    names, tab = prof.counter_table([str(p)])
    assert names == ["TCC_HIT_sum", "TCC_MISS_sum"]
    assert tab["k_pred_rows"]["TCC_HIT_sum"] == (2, 80.0)
    prof.main(["counters", str(p)])
    out = capsys.readouterr().out
    assert "| k_pred_rows | pred | 2 | 80 | 20 | 80.0 |" in out


def write_trace(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Stream_Id", "Grid_Size_Z", "Workgroup_Size_Z",
                    "Start_Timestamp", "End_Timestamp"])
        w.writerows(rows)


def test_timeline_window_gaps_and_classes(tmp_path, capsys):
    """Window between the last two markers; kernels of two classes overlap,
    one gap of 10 ns; a dispatch outside the window is ignored."""
    p = tmp_path / "t.csv"
    write_trace(p, [(COPY, 0, 1, 1, 0, 5),                 # earlier marker, ignored
                    (PRED, 1, 19, 1, 6, 50),               # before the window
                    (COPY, 0, 1, 1, 100, 110),             # marker: window opens at 110
                    (PRED, 1, 19, 1, 110, 150),
                    (PRED, 2, 15, 1, 120, 140),
                    (PRED, 1, 19, 1, 160, 190),
                    (COPY, 0, 1, 1, 200, 205)])            # marker: window closes at 200
    t = prof.timeline(str(p))
    assert t["window_ns"] == 90 and t["dispatches"] == 3
    assert t["busy_ns"] == 70 and t["kernel_ns"] == 90
    (a, b) = t["classes"]
    assert (a["members"], a["kernel_ns"], a["busy_ns"]) == (19, 70, 70)
    assert (b["members"], b["kernel_ns"]) == (15, 20)
    prof.main(["timeline", str(p), "--steps", "1"])
    out = capsys.readouterr().out
    assert "no kernel running" in out and "(22.2%)" in out


def test_union_length():
    assert prof.union_length([(0, 10), (5, 15), (20, 30)]) == 25
    assert prof.union_length([]) == 0
