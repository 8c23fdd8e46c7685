"""CPU checks of the HBM traffic summary (cosmictiger_amd/traffic.py) that the
bench's roofline `traffic` comes from: counter CSVs in rocprofv3's layout
with known values give the closed-form bytes, the calibration scales each
counter, kicks are averaged, and the output file keeps one entry per bucket."""
import csv
import json
import os

import pytest


def _write_run(d, counter, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for i, (name, v) in enumerate(rows):
            # a counter can be reported in several rows (per XCD); they add up
            w.writerow({"Dispatch_Id": i, "Kernel_Name": name, "Counter_Name": counter, "Counter_Value": v / 2})
            w.writerow({"Dispatch_Id": i, "Kernel_Name": name, "Counter_Name": counter, "Counter_Value": v / 2})


def _runs(tmp_path, pp_fetch, pp_write):
    copy_mib = 64
    # the copy reads 64 MiB but FETCH_SIZE reports half of it (scale 2),
    # WRITE_SIZE reports all of it (scale 1); two kicks, two P2P classes each
    fetch = [("__amd_rocclr_copyBuffer", copy_mib * 1024 / 2)]
    write = [("__amd_rocclr_copyBuffer", copy_mib * 1024)]
    for _ in range(2):
        fetch += [("void ct::k_pp<4, 2, 0>(ct::PPArgs)", pp_fetch[0]), ("void ct::k_pp<5, 4, 0>(ct::PPArgs)", pp_fetch[1]),
                  ("void ct::k_kick<0>(ct::LeafArgs)", 1e6)]
        write += [("void ct::k_pp<4, 2, 0>(ct::PPArgs)", pp_write[0]), ("void ct::k_pp<5, 4, 0>(ct::PPArgs)", pp_write[1]),
                  ("void ct::k_kick<0>(ct::LeafArgs)", 1e6)]
    _write_run(str(tmp_path / "f"), "FETCH_SIZE", fetch)
    _write_run(str(tmp_path / "w"), "WRITE_SIZE", write)
    return copy_mib


def test_summary_closed_form(tmp_path):
    from cosmictiger_amd import traffic
    copy_mib = _runs(tmp_path, (1000.0, 3000.0), (10.0, 30.0))
    res = traffic.summarize(str(tmp_path / "f"), str(tmp_path / "w"), copy_mib)
    assert res["fetch"]["scale"] == pytest.approx(2.0)
    assert res["write"]["scale"] == pytest.approx(1.0)
    assert res["fetch"]["kicks"] == 2
    # per kick: (1000 + 3000) KiB raw, times 1024 bytes, times the scale
    assert res["fetch"]["pp_bytes_per_kick"] == pytest.approx(4000 * 1024 * 2.0)
    assert res["write"]["pp_bytes_per_kick"] == pytest.approx(40 * 1024 * 1.0)


def test_output_keeps_one_entry_per_bucket(tmp_path):
    from cosmictiger_amd import traffic
    copy_mib = _runs(tmp_path, (1000.0, 3000.0), (10.0, 30.0))
    out = str(tmp_path / "t.json")
    common = [str(tmp_path / "f"), str(tmp_path / "w"), "--copy-mib", str(copy_mib), "--n", "4096", "--workload", "clustered", "--out", out]
    assert traffic.main(common + ["--bucket", "64"]) == 0
    assert traffic.main(common + ["--bucket", "160"]) == 0
    assert traffic.main(common + ["--bucket", "64"]) == 0  # a new measurement replaces the old one
    entries = json.load(open(out))["entries"]
    assert sorted(e["bucket"] for e in entries) == [64, 160]
    for e in entries:
        assert e["bytes_per_launch"] == pytest.approx(4000 * 1024 * 2.0 + 40 * 1024)
