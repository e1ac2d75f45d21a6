"""Per-pass summary of rocprofv3 runs of bench.py's 512^3 workload.

Three runs of the same bench command give, per kernel, the average duration
(``--kernel-trace --stats``) and the HBM bytes per launch (``--pmc FETCH_SIZE``
and ``--pmc WRITE_SIZE``, in KiB per dispatch).  The counters are calibrated
against the runtime's device-to-device copy of the 513^3 right-hand side
(``__amd_rocclr_copyBuffer``, 513^3 doubles read and written): the factor
that turns its counter value into its known byte count is applied to every
kernel.  The result, one line per pass of the solves and the transform, is
what ``bench.py`` reads for ``roofline.traffic``:

    python -m cuhelmholtz_amd.profsum --stats S --fetch F --write W \\
        [--out profiles/traffic_512.txt]

S, F and W are the ``-d`` output folders of the three rocprofv3 runs (CSV
output).  The passes are matched to kernels by operation and kind selector
(the kernel's last template argument) for the workload's boundary codes
(3, 4, 0) and transform kinds (DCT-I, DCT-III, R2C); two pairs of passes run
the same kernel (the solve's x inverse and the transform's x pass; the
solve's y forward pass and the transform's y pass), and their lines then
carry that kernel's averages.
"""
import argparse
import csv
import glob
import os
import re
import statistics
import sys

from . import kernelgen as G

# pass -> (precision, op, kind selector) for codes (3, 4, 0) and transform kinds (0, 2, 6)
BC = (3, 4, 0)
PASSES = [
    ("xfwd", 64, G.B_XFWD, G.FWD_KINDS[BC[0]]),
    ("yfwd", 64, G.B_COLS, G.FWD_KINDS[BC[1]]),
    ("zmid", 64, G.B_ZMID, BC[2]),
    ("yinv", 64, G.B_COLS, G.INV_KINDS[BC[1]]),
    ("xinv", 64, G.B_ROWS, G.INV_KINDS[BC[0]]),
    ("tx", 64, G.B_ROWS, G.K_DCT_1),
    ("ty", 64, G.B_COLS, G.K_DCT_3),
    ("tz", 64, G.B_COLS, G.K_R2C),
    ("mixed:xfwd", 32, G.B_XFWD, G.FWD_KINDS[BC[0]]),
    ("mixed:yfwd", 32, G.B_COLS, G.FWD_KINDS[BC[1]]),
    ("mixed:zmid", 32, G.B_ZMID, BC[2]),
    ("mixed:yinv", 32, G.B_COLS, G.INV_KINDS[BC[1]]),
    ("mixed:xinv/xupd", 32, G.B_XUPD, G.INV_KINDS[BC[0]]),
    ("mixed:xres", 32, G.B_XRES, G.FWD_KINDS[BC[0]]),
]
COPY = "__amd_rocclr_copyBuffer"
COPY_BYTES = 513 ** 3 * 8
NAME = re.compile(r"k_line<(double|float), (\d+), 512, (\d+), (\d+), (\w+), (\d+)(?:, (-?\d+))?>")


def _csv(folder, suffix):
    hits = sorted(glob.glob(os.path.join(folder, "**", "*" + suffix), recursive=True))
    if not hits:
        raise SystemExit(f"no *{suffix} under {folder}")
    return hits[0]


def _counter(folder, name):
    per = {}
    for r in csv.DictReader(open(_csv(folder, "counter_collection.csv"))):
        if r["Counter_Name"] == name:
            per.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return per


def _key(kernel):
    """(precision, op, kind selector) of a k_line kernel name, or None."""
    m = NAME.search(kernel)
    if not m:
        return None
    return (64 if m.group(1) == "double" else 32, int(m.group(2)), int(m.group(7) or -1))


def summarise(stats_dir, fetch_dir, write_dir):
    """Rows (pass, read GB, write GB, average us, launches, kernel)."""
    avg, calls = {}, {}
    for r in csv.DictReader(open(_csv(stats_dir, "kernel_stats.csv"))):
        avg[r["Name"]] = float(r["AverageNs"]) / 1e3
        calls[r["Name"]] = int(r["Calls"])
    fetch, write = _counter(fetch_dir, "FETCH_SIZE"), _counter(write_dir, "WRITE_SIZE")
    if COPY not in fetch or COPY not in write:
        raise SystemExit(f"no {COPY} dispatch to calibrate against")
    fs = COPY_BYTES / (max(fetch[COPY]) * 1024)
    ws = COPY_BYTES / (max(write[COPY]) * 1024)
    by_key = {}
    for k in fetch:
        key = _key(k)
        if key is not None:
            by_key.setdefault(key, []).append(k)
    rows = []
    for name, prec, op, ks in PASSES:
        ks_ = by_key.get((prec, op, ks)) or by_key.get((prec, op, -1)) or []
        if len(ks_) != 1:
            continue
        k = ks_[0]
        rd = statistics.median(fetch[k]) * 1024 * fs / 1e9
        wr = statistics.median(write.get(k, [0.0])) * 1024 * ws / 1e9
        short = k.split("(")[0].replace("void ", "")
        rows.append((name, rd, wr, avg.get(k, float("nan")), calls.get(k, 0), short))
    return rows, fs, ws


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--stats", required=True, help="rocprofv3 --kernel-trace --stats output folder")
    ap.add_argument("--fetch", required=True, help="rocprofv3 --pmc FETCH_SIZE output folder")
    ap.add_argument("--write", required=True, help="rocprofv3 --pmc WRITE_SIZE output folder")
    ap.add_argument("--out", default=None, help="traffic table to write (default: stdout)")
    a = ap.parse_args(argv)
    rows, fs, ws = summarise(a.stats, a.fetch, a.write)
    lines = ["# HBM bytes per launch of the kernel behind each pass of the 512^3 solves and 3D transform",
             "# (FETCH_SIZE x %.3f + WRITE_SIZE x %.3f, calibrated on %s of 513^3 doubles), and" % (fs, ws, COPY),
             "# the kernel's average duration under rocprofv3 --kernel-trace --stats; written by",
             "# python -m cuhelmholtz_amd.profsum; FP64 passes by name, transform passes tx/ty/tz,",
             "# mixed-precision passes as mixed:<name>",
             "# pass              read_GB  write_GB   avg_us  launches  kernel"]
    for name, rd, wr, us, n, k in rows:
        lines.append(f"{name:18s} {rd:8.3f} {wr:9.3f} {us:8.1f} {n:9d}  {k}")
    text = "\n".join(lines) + "\n"
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    else:
        sys.stdout.write(text)


if __name__ == "__main__":
    main()
