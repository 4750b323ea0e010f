"""The bench's live HBM traffic reading, on the CPU: parsing rocprofv3's counter CSV and scaling the
roofline kernel's FETCH_SIZE / WRITE_SIZE by the same counter's reading for the 1 GiB copies of the
same pass (bench.pmc_dispatches, bench.live_traffic), with the profiler replaced by a stand-in that
writes the CSV a pass would."""
import csv
import os

import pytest

import bench

FUSED = "void cg::k_relax_tile<4, 6, false, 3>(cg::LevelArgs, float const*, float const*, float*, cg::TileExtra)"
COPY = "__amd_rocclr_copyBuffer"


def write_counters(path, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for d, k, c, v in rows:
            w.writerow({"Dispatch_Id": d, "Kernel_Name": k, "Counter_Name": c, "Counter_Value": v})


def test_counter_rows_of_one_dispatch_are_summed_and_other_counters_ignored(tmp_path):
    write_counters(str(tmp_path / "host" / "123" / "run_counter_collection.csv"),
                   [(1, FUSED, "FETCH_SIZE", 10.0), (1, FUSED, "FETCH_SIZE", 5.0), (2, FUSED, "FETCH_SIZE", 14.0),
                    (3, COPY, "FETCH_SIZE", 7.0), (3, COPY, "WRITE_SIZE", 99.0)])
    got = bench.pmc_dispatches(str(tmp_path), "FETCH_SIZE")
    assert sorted(got[FUSED]) == [14.0, 15.0]
    assert got[COPY] == [7.0]
    assert bench.pmc_dispatches(str(tmp_path), "WRITE_SIZE") == {COPY: [99.0]}


def test_no_profiler_reports_why(monkeypatch):
    monkeypatch.setattr("shutil.which", lambda name: None)
    assert bench.live_traffic(6, "6,6,4,4,4") == (None, None, "rocprofv3 not found")


class FakePass:
    """Stands in for one rocprofv3 pass: writes the counter CSV into the pass's -d directory."""
    readings = {"FETCH_SIZE": {"copies": [524300.0, 524301.0, 524299.0, 12.0, 3.0], "fused": [23700.0, 23710.0, 23690.0]},
                "WRITE_SIZE": {"copies": [1048576.0, 1048576.0, 1048576.0, 8.0], "fused": [24659.0, 24660.0]}}
    status = 0

    def __init__(self, cmd, **kw):
        counter = cmd[cmd.index("--pmc") + 1]
        out = cmd[cmd.index("-d") + 1]
        r = self.readings[counter]
        rows = [(i, COPY, counter, v) for i, v in enumerate(r["copies"])]
        rows += [(100 + i, FUSED, counter, v) for i, v in enumerate(r["fused"])]
        write_counters(os.path.join(out, "host", "1", "run_counter_collection.csv"), rows)
        self.pid = os.getpid()

    def wait(self, timeout=None):
        return self.status


def test_traffic_is_scaled_by_the_three_largest_copies(monkeypatch):
    monkeypatch.setattr("shutil.which", lambda name: "/usr/bin/rocprofv3")
    monkeypatch.setattr(bench.subprocess, "Popen", FakePass)
    read, written, note = bench.live_traffic(6, "6,6,4,4,4")
    # Reads: median of the fused dispatches over the median of the three largest copy dispatches (the
    # plan's small uploads go through the same copy kernel and are left out), times 1 GiB.
    assert read == pytest.approx(23700.0 / 524300.0 * (1 << 30))
    assert written == pytest.approx(24659.5 / 1048576.0 * (1 << 30))
    assert "1 GiB" in note


def test_a_failed_pass_falls_back(monkeypatch):
    monkeypatch.setattr("shutil.which", lambda name: "/usr/bin/rocprofv3")

    class Failing(FakePass):
        status = 1

    monkeypatch.setattr(bench.subprocess, "Popen", Failing)
    read, written, note = bench.live_traffic(6, "6,6,4,4,4")
    assert read is None and written is None and "no reading" in note
