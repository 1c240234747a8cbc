"""The counter summary behind roofline.traffic: per-dispatch sums, the copy
scale, and per-kernel averages (host logic only)."""
import csv

import pytest

from partmc_amd import pmc_traffic

FIELDS = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def write_csv(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(FIELDS, r)))
    return str(path)


def test_summary_scales_by_copy(tmp_path):
    nbytes = 1 << 20
    # the copy reads as half its size in FETCH_SIZE units and exactly in WRITE_SIZE
    cf = write_csv(tmp_path / "cf.csv", [(1, "fill(void*)", "FETCH_SIZE", 0.0),
                                         (2, "copy(void*)", "FETCH_SIZE", nbytes / 2048)])
    cw = write_csv(tmp_path / "cw.csv", [(1, "fill(void*)", "WRITE_SIZE", nbytes / 1024),
                                         (2, "copy(void*)", "WRITE_SIZE", nbytes / 1024)])
    # a kernel with two dispatches, one of them split over two counter rows
    f = write_csv(tmp_path / "f.csv", [(7, "pmc::k_a(int)", "FETCH_SIZE", 10.0),
                                       (8, "pmc::k_a(int)", "FETCH_SIZE", 20.0),
                                       (8, "pmc::k_a(int)", "FETCH_SIZE", 10.0),
                                       (9, "pmc::k_b()", "OTHER", 5.0)])
    w = write_csv(tmp_path / "w.csv", [(7, "pmc::k_a(int)", "WRITE_SIZE", 4.0),
                                       (8, "pmc::k_a(int)", "WRITE_SIZE", 6.0)])
    s = pmc_traffic.summarize(cf, cw, f, w, nbytes)
    assert s["fetch_scale"] == pytest.approx(2048.0)
    assert s["write_scale"] == pytest.approx(1024.0)
    k = s["kernels"]["k_a"]
    assert k["dispatches"] == 2
    assert k["fetch_bytes"] == pytest.approx(2048.0 * (10.0 + 30.0) / 2)
    assert k["write_bytes"] == pytest.approx(1024.0 * (4.0 + 6.0) / 2)
    assert k["bytes"] == pytest.approx(k["fetch_bytes"] + k["write_bytes"])
    assert "k_b" not in s["kernels"]
