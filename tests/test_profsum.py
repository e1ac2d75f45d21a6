"""CPU checks of the rocprofv3 summary behind profiles/traffic_512.txt: the
copy-kernel calibration, the pass -> kernel matching by operation and kind
selector, and the table format bench.py reads back."""
import csv
import os

import pytest

from cuhelmholtz_amd import profsum as P

COPY_KIB_FETCH = P.COPY_BYTES / 1024 / 2.0   # the counter reads half the bytes (x2 calibration)
COPY_KIB_WRITE = P.COPY_BYTES / 1024


def _write(folder, name, header, rows):
    os.makedirs(folder, exist_ok=True)
    with open(os.path.join(folder, name), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def _fixture(tmp_path, kernels):
    """kernels: {name: (avg_ns, calls, fetch KiB, write KiB)} plus the copy."""
    s, fe, wr = (str(tmp_path / d) for d in ("s", "f", "w"))
    _write(s, "run_kernel_stats.csv", ["Name", "Calls", "AverageNs"],
           [[k, v[1], v[0]] for k, v in kernels.items()])
    frows = [[1, P.COPY, "FETCH_SIZE", COPY_KIB_FETCH]]
    wrows = [[1, P.COPY, "WRITE_SIZE", COPY_KIB_WRITE]]
    for i, (k, v) in enumerate(kernels.items()):
        for d in range(3):
            frows.append([10 + 3 * i + d, k, "FETCH_SIZE", v[2]])
            wrows.append([10 + 3 * i + d, k, "WRITE_SIZE", v[3]])
    hdr = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    _write(fe, "run_counter_collection.csv", hdr, frows)
    _write(wr, "run_counter_collection.csv", hdr, wrows)
    return s, fe, wr


def test_calibration_and_matching(tmp_path):
    gib = 2 ** 30
    ks = {
        # FP64 fused z pass, code 0 (periodic)
        "chz::k_line<double, 3, 512, 16, 16, 4112ull, 16, 0>(chz::PassArgs, chz::SolveArgs, double const*, double*)":
            (560000.0, 13, gib / 1024 / 2.0, gib / 1024),
        # FP64 y pass with DCT-III (the solve's y forward pass and the transform's y pass)
        "chz::k_line<double, 1, 512, 32, 16, 264200ull, 8, 2>(chz::PassArgs, chz::SolveArgs, double const*, double*)":
            (480000.0, 19, 1.1 * gib / 1024 / 2.0, gib / 1024),
        # a kernel for every kind, not matched where a kind-specialised one exists
        "chz::k_line<double, 1, 512, 32, 16, 264200ull, 8>(chz::PassArgs, chz::SolveArgs, double const*, double*)":
            (999000.0, 1, 9.0, 9.0),
    }
    rows, fs, ws = P.summarise(*_fixture(tmp_path, ks))
    assert fs == pytest.approx(2.0) and ws == pytest.approx(1.0)
    got = {r[0]: r for r in rows}
    assert got["zmid"][1] == pytest.approx(gib / 1e9) and got["zmid"][2] == pytest.approx(gib / 1e9)
    assert got["zmid"][3] == pytest.approx(560.0) and got["zmid"][4] == 13
    assert got["yfwd"][1] == pytest.approx(1.1 * gib / 1e9)
    assert got["ty"][5] == got["yfwd"][5]          # the shared kernel
    # no kernel of its kind in the profile: the pass ran on the kernel for every kind
    assert got["yinv"][5] == "chz::k_line<double, 1, 512, 32, 16, 264200ull, 8>"
    assert "xfwd" not in got                        # no kernel of its operation at all


def test_table_read_by_bench(tmp_path, monkeypatch):
    gib = 2 ** 30
    ks = {"chz::k_line<double, 3, 512, 16, 16, 4112ull, 16, 0>(args)": (560000.0, 13, gib / 2048.0, gib / 1024)}
    s, fe, wr = _fixture(tmp_path, ks)
    out = tmp_path / "profiles" / "traffic_512.txt"
    out.parent.mkdir()
    P.main(["--stats", s, "--fetch", fe, "--write", wr, "--out", str(out)])
    import bench
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    traffic, kernel = bench.measured_traffic("zmid")
    assert traffic == pytest.approx(2 * gib, rel=1e-3)
    assert kernel == "chz::k_line<double, 3, 512, 16, 16, 4112ull, 16, 0>"
    assert bench.measured_traffic("xfwd") == (None, None)
