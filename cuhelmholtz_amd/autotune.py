"""Autotuner of the pass launch configurations (elements per thread, threads
per line, lines per workgroup -> radix plan) for each kernel role.

For every candidate the pass is first checked against the output of the
built-in configuration on the same input (correctness does not depend on the
choice; a candidate that differs by more than the precision's tolerance is
rejected), then timed with the library's per-pass events.  The winners are
written as a tuning table ("role N prec ept tpl lines" per line) that the
library loads at start-up from tuning/<gfx arch>.txt beside the shared object.

Usage (on the GPU):
  python -m cuhelmholtz_amd.autotune --solve 512 512 512 --prec 64 [--out FILE]
  python -m cuhelmholtz_amd.autotune --transform 256 --prec 64
  python -m cuhelmholtz_amd.autotune --batch 32 48 64 96 128 --prec 32
"""
import argparse
import os
import sys

import cuhelmholtz_amd as C

HERE = os.path.dirname(os.path.abspath(__file__))
EPTS = (16, 12, 10, 8, 6, 5, 4)
LINES = (1, 2, 4, 6, 8, 10, 12, 14, 16, 20, 24, 28, 32, 48, 64, 96, 128, 256, 512)


def candidates(N):
    M = N // 2
    out = []
    for ept in EPTS:
        base = -(-M // ept)
        for tpl in sorted({base, base + 1, 2 * base}):
            if tpl > 256:
                continue
            for lines in LINES:
                if tpl * lines > 512 or tpl * lines < 64:
                    continue
                out.append((ept, tpl, lines))
    return out


def _median(v):
    v = sorted(v)
    return v[len(v) // 2]


def _pass_ms(run, names, reps):
    """median over reps of the summed duration of the named passes"""
    import torch
    run()
    torch.cuda.synchronize()
    C.set_timing(True)
    for _ in range(reps):
        run()
    torch.cuda.synchronize()
    rec = C.get_timing(64 * reps)
    C.set_timing(False)
    per_call, cur = [], 0.0
    seen = 0
    for name, ms in rec:
        if name in names:
            cur += ms
            seen += 1
            if seen == len(names):
                per_call.append(cur)
                cur, seen = 0.0, 0
    return _median(per_call) if per_call else float("inf")


def tune_role(N, prec, role, run, check, names, reps=5, log=print):
    """Try every candidate of (N, prec, role); returns (best config, ms, table)."""
    results = []
    for ept, tpl, lines in candidates(N):
        try:
            C.set_axis_config(N, prec, role, ept, tpl, lines)
        except C.CHZError:
            continue
        if not check():
            log(f"  reject {ept} {tpl} {lines}: result differs")
            continue
        ms = _pass_ms(run, names, reps)
        results.append((ms, ept, tpl, lines))
    results.sort()
    best = results[0]
    C.set_axis_config(N, prec, role, best[1], best[2], best[3])
    return best, results


def tune_solve(n, prec=64, bc=(3, 4, 0), lam=-4.0, log=print):
    import torch
    from cuhelmholtz_amd import workloads as W
    dev = torch.device("cuda:0")
    f, bcs = W.smooth_modes(n)
    f0 = torch.from_numpy(f).to(dev)
    bct = [None if b is None else torch.from_numpy(b).to(dev) for b in bcs]
    del f
    solve = C.helmholtz_solve if prec == 64 else C.helmholtz_solve_mixed_fp32only
    C.reset_axis_configs()
    ref = f0.clone()
    solve(ref, n, bc, lam, bct)
    work = f0.clone()
    tol = 1e-12 if prec == 64 else 1e-5

    def run():
        work.copy_(f0)
        solve(work, n, bc, lam, bct)

    def check():
        run()
        torch.cuda.synchronize()
        return ((work - ref).abs().max() / ref.abs().max()).item() < tol

    table = []
    plan = [(C.ROLE_COLS, n[1], ("yfwd", "yinv")), (C.ROLE_ZMID, n[2], ("zmid",)),
            (C.ROLE_XFWD, n[0], ("xfwd",)), (C.ROLE_ROWS, n[0], ("xinv",))]
    for role, N, names in plan:
        best, res = tune_role(N, prec, role, run, check, names, log=log)
        log(f"role {C.ROLE_NAMES[role]} N={N} prec={prec}: best ept={best[1]} tpl={best[2]} lines={best[3]} "
            f"{best[0]:.3f} ms ({len(res)} candidates; worst {res[-1][0]:.3f} ms)")
        table.append((role, N, prec, best[1], best[2], best[3], best[0], res))
    return table


def tune_transform(N, prec=64, log=print):
    import torch
    dev = torch.device("cuda:0")
    dt = torch.float64 if prec == 64 else torch.float32
    n = (N, N, N)
    x = torch.rand((N + 2, N + 2, N + 2), dtype=dt, device=dev)
    kinds = (C.DCT_2, C.DCT_2, C.DCT_2)
    C.reset_axis_configs()
    ref = torch.empty_like(x)
    C.do_transform(x, ref, kinds)
    y = torch.empty_like(x)
    tol = 1e-12 if prec == 64 else 1e-5

    def run():
        C.do_transform(x, y, kinds)

    def check():
        run()
        torch.cuda.synchronize()
        sl = (slice(0, N), slice(0, N), slice(0, N))
        return ((y[sl] - ref[sl]).abs().max() / ref[sl].abs().max()).item() < tol

    table = []
    for role, names in [(C.ROLE_ROWS, ("x",)), (C.ROLE_COLS, ("y",)), (C.ROLE_ZCOLS, ("z",))]:
        best, res = tune_role(N, prec, role, run, check, names, log=log)
        log(f"role {C.ROLE_NAMES[role]} N={N} prec={prec}: best ept={best[1]} tpl={best[2]} lines={best[3]} "
            f"{best[0]:.3f} ms ({len(res)} candidates)")
        table.append((role, N, prec, best[1], best[2], best[3], best[0], res))
    return table


def tune_batch(N, prec=64, count=32, reps=3, log=print):
    """Tune the roles of one size class of the batched solve: a batch of
    `count` N^3 problems (codes and lambda drawn like the batched workload),
    timed per launch through direct (non-graph) batched solves."""
    import numpy as np
    import torch
    from cuhelmholtz_amd import workloads as W
    dev = torch.device("cuda:0")
    descs = W.batched_problems(count=count, seed=7, sizes=(N,))
    probs, f0 = [], []
    for d in descs:
        f, bcs = W.problem_arrays(d)
        ft = torch.from_numpy(f).to(dev)
        f0.append(ft.clone())
        probs.append((ft, d["n"], d["bc"], d["lam"], [None if b is None else torch.from_numpy(b).to(dev) for b in bcs]))
    refine = 0 if prec == 64 else 1
    C.reset_axis_configs()

    def solve_into():
        for p, g in zip(probs, f0):
            p[0].copy_(g)
        bs = C.BatchSolver(probs, prec, refine)
        bs.solve(graph=False)
        torch.cuda.synchronize()
        return bs

    solve_into().close()
    ref = [p[0].clone() for p in probs]
    tol = 1e-12 if prec == 64 else 1e-9

    def check():
        solve_into().close()
        return all(float((p[0] - r).abs().max() / r.abs().max()) < tol for p, r in zip(probs, ref))

    def timed(names):
        for p, g in zip(probs, f0):
            p[0].copy_(g)
        bs = C.BatchSolver(probs, prec, refine)
        bs.solve(graph=False)
        torch.cuda.synchronize()
        out = []
        for _ in range(reps):
            C.set_timing(True)
            bs.solve(graph=False)
            torch.cuda.synchronize()
            rec = C.get_timing(4096)
            C.set_timing(False)
            out.append(sum(ms for nm, ms in rec if nm.split("/")[0] in names))
        bs.close()
        return float(np.median(out))

    plan = [(C.ROLE_COLS, ("yfwd", "yinv")), (C.ROLE_ZMID, ("zmid",)), (C.ROLE_XFWD, ("xfwd", "xres")),
            (C.ROLE_ROWS, ("xinv", "xupd"))]
    table = []
    for role, names in plan:
        results = []
        for ept, tpl, lines in candidates(N):
            try:
                C.set_axis_config(N, prec, role, ept, tpl, lines)
            except C.CHZError:
                continue
            if not check():
                log(f"  reject {ept} {tpl} {lines}: result differs")
                continue
            results.append((timed(names), ept, tpl, lines))
        results.sort()
        best = results[0]
        C.set_axis_config(N, prec, role, best[1], best[2], best[3])
        log(f"batch role {C.ROLE_NAMES[role]} N={N} prec={prec}: best ept={best[1]} tpl={best[2]} lines={best[3]} "
            f"{best[0]:.3f} ms ({len(results)} candidates; worst {results[-1][0]:.3f} ms)")
        table.append((role, N, prec, best[1], best[2], best[3], best[0], results))
    return table


def write_table(path, rows, header):
    """merge rows into a tuning table file (later rows win per (role, N, prec))"""
    entries = {}
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                if line.strip() and not line.startswith("#"):
                    v = line.split()
                    entries[(int(v[0]), int(v[1]), int(v[2]))] = line.rstrip("\n")
    for r in rows:
        role, N, prec, ept, tpl, lines, ms = r[:7]
        entries[(role, N, prec)] = f"{role} {N} {prec} {ept} {tpl} {lines}   # {C.ROLE_NAMES[role]} {ms:.3f} ms"
    with open(path, "w") as f:
        f.write(header)
        for k in sorted(entries):
            f.write(entries[k] + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--solve", type=int, nargs=3)
    ap.add_argument("--transform", type=int)
    ap.add_argument("--batch", type=int, nargs="*", help="size classes of the batched solve to tune")
    ap.add_argument("--prec", type=int, default=64)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning", "gfx950.txt"))
    args = ap.parse_args()
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    rows = []
    if args.solve:
        rows += tune_solve(tuple(args.solve), args.prec)
    if args.transform:
        rows += tune_transform(args.transform, args.prec)
    for N in args.batch or []:
        rows += tune_batch(N, args.prec)
    header = ("# launch configurations chosen by cuhelmholtz_amd.autotune on MI355X (gfx950)\n"
              "# role N prec ept tpl lines   (role: 0 rows, 1 cols, 2 solve xfwd, 3 solve zmid, 4 z cols)\n")
    write_table(args.out, rows, header)
    for r in rows:
        print(f"{C.ROLE_NAMES[r[0]]} N={r[1]} prec={r[2]}: top 5 " +
              "; ".join(f"{ms:.3f}ms e{e} t{t} l{l}" for ms, e, t, l in r[7][:5]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
