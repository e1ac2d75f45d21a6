"""Autotuner of the pass launch configurations (elements per thread, threads
per line, lines per workgroup -> radix plan) for each kernel role.

The tuner owns the compiled kernel shapes: the fixed-shape kernels are built
only for the rows of its two tables (cuhelmholtz_amd/kernelgen.py generates
their instantiations at build time), so tuning is a loop of

  1. `--emit-alt K`: time every legal candidate (shapes without a compiled
     kernel run on the generic kernels) and append the K fastest per role to
     tuning/<arch>.alt.txt, the shapes compiled beside the chosen ones;
  2. rebuild (python -m cuhelmholtz_amd.build) so those shapes get their
     fixed-shape kernels;
  3. `--compiled-only`: time the compiled candidates and write the winners
     into tuning/<arch>.txt ("role N prec ept tpl lines" per line), which the
     library loads at start-up (and the next build compiles).

Every candidate is checked before it is timed, by properties that hold for
the right answer independently of any launch configuration, plus agreement
with the built-in configuration's result:
  * solves: the normalised 7-point residual |f - A u| / (|A| |u| + |f|) of
    the discrete system (cuhelmholtz_amd/checks.py) below 1e-13 (FP64) or
    2e-5 (FP32 without refinement: about 3e-6 at 512^3);
  * transforms: the round trip DCT-III(DCT-II(x)) = (N/2)^3 x per axis
    triple, to 1e-12 (FP64) or 1e-5 (FP32);
  * batched solves: the residual of every problem of the batch.
A candidate failing a check is rejected and reported.  (The tests check
tuned and built-in configurations against the CPU oracle.)

Usage (on the GPU):
  python -m cuhelmholtz_amd.autotune --solve 512 512 512 --prec 64 [--out FILE]
  python -m cuhelmholtz_amd.autotune --transform 256 --prec 64
  python -m cuhelmholtz_amd.autotune --batch 32 48 64 96 128 --prec 32
  ... [--emit-alt K] [--compiled-only]
"""
import argparse
import os
import sys

import cuhelmholtz_amd as C
from cuhelmholtz_amd import kernelgen as G

HERE = os.path.dirname(os.path.abspath(__file__))
EPTS = G.EPTS
LINES = (1, 2, 4, 6, 8, 10, 12, 14, 16, 20, 24, 28, 32, 48, 64, 96, 128, 256, 512)
RES_TOL = {64: 1e-13, 32: 2e-5}
AGREE_TOL = {64: 1e-12, 32: 1e-5}


def candidates(N):
    """Legal (ept, tpl, lines) shapes of an N-cell axis: per points-per-thread
    value, the fewest threads per line that fit, one more, and twice as many,
    with 64 to 512 threads per tile."""
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


def compiled_shapes(N, prec, role):
    """(ept, tpl, lines) of N with a compiled single-problem kernel for every
    pass of the role."""
    have = {}
    for form, p, op, n, ept, tpl, lines, plan, ks in C.compiled_kernels():
        if p == prec and n == N and form in (G.F_LINE, G.F_FIXED) and ks < 0:
            have.setdefault((ept, tpl, lines), set()).add(op)
    need = set(G.role_ops(role, prec))
    return sorted(k for k, ops in have.items() if need <= ops)


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


def tune_role(N, prec, role, run, check, names, reps=5, log=print, shapes=None):
    """Check and time every candidate shape of (N, prec, role) (shapes: a
    subset to try); pins the fastest and returns (best, results) with results
    sorted (ms, ept, tpl, lines)."""
    results = []
    for ept, tpl, lines in shapes if shapes is not None else candidates(N):
        try:
            C.set_axis_config(N, prec, role, ept, tpl, lines)
        except C.CHZError:
            continue
        ok, why = check()
        if not ok:
            log(f"  reject ept={ept} tpl={tpl} lines={lines}: {why}")
            continue
        ms = _pass_ms(run, names, reps)
        results.append((ms, ept, tpl, lines))
    if not results:
        raise C.CHZError(f"no candidate of role {C.ROLE_NAMES[role]} N={N} prec={prec} passed its checks")
    results.sort()
    best = results[0]
    C.set_axis_config(N, prec, role, best[1], best[2], best[3])
    return best, results


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


def tune_solve(n, prec=64, bc=(3, 4, 0), lam=-4.0, log=print, compiled_only=False, roles=None, reps=5):
    """Tune the four roles of a solve of size n (FP64, or the FP32 solve the
    mixed path refines).  Returns rows (role, N, prec, ept, tpl, lines, ms, results)."""
    import torch
    from cuhelmholtz_amd import checks
    from cuhelmholtz_amd import workloads as W
    dev = torch.device("cuda:0")
    # the benchmark rhs for its codes, else N(0,1) data with the face arrays the codes need
    f, bcs = W.smooth_modes(n) if tuple(bc) == (3, 4, 0) else W.problem_arrays({"n": n, "bc": bc, "seed": 5})
    f0 = torch.from_numpy(f).to(dev)
    bct = [None if b is None else torch.from_numpy(b).to(dev) for b in bcs]
    del f
    solve = C.helmholtz_solve if prec == 64 else C.helmholtz_solve_mixed_fp32only
    C.reset_axis_configs()
    ref = f0.clone()
    solve(ref, n, bc, lam, bct)
    work = f0.clone()

    def run():
        work.copy_(f0)
        solve(work, n, bc, lam, bct)

    def check():
        run()
        torch.cuda.synchronize()
        res = checks.relative_residual(work, f0, n, bc, lam, bcs)
        if not res < RES_TOL[prec]:
            return False, f"relative residual {res:.2e}"
        d = _rel(work, ref)
        if not d < AGREE_TOL[prec]:
            return False, f"differs from the built-in configuration by {d:.2e}"
        return True, ""

    table = []
    # the fused z pass has a role per z periodicity (its transforms differ)
    zrole = C.ROLE_ZMID if bc[2] == 0 else C.ROLE_ZMIDN
    plan = [(C.ROLE_COLS, n[1], ("yfwd", "yinv")), (zrole, n[2], ("zmid",)),
            (C.ROLE_XFWD, n[0], ("xfwd",)), (C.ROLE_ROWS, n[0], ("xinv",))]
    for role, N, names in plan:
        if roles is not None and role not in roles:
            continue
        shapes = compiled_shapes(N, prec, role) if compiled_only else None
        best, res = tune_role(N, prec, role, run, check, names, log=log, shapes=shapes, reps=reps)
        log(f"role {C.ROLE_NAMES[role]} N={N} prec={prec}: best ept={best[1]} tpl={best[2]} lines={best[3]} "
            f"{best[0]:.3f} ms ({len(res)} candidates; worst {res[-1][0]:.3f} ms)")
        table.append((role, N, prec, best[1], best[2], best[3], best[0], res))
    return table


def tune_transform(N, prec=64, log=print, compiled_only=False, reps=5):
    """Tune the three roles of an N^3 DCT-II transform (x rows, y columns, z
    columns)."""
    import torch
    dev = torch.device("cuda:0")
    dt = torch.float64 if prec == 64 else torch.float32
    x = torch.rand((N + 2, N + 2, N + 2), dtype=dt, device=dev) * 2 - 1
    kinds = (C.DCT_2, C.DCT_2, C.DCT_2)
    inv = (C.DCT_3, C.DCT_3, C.DCT_3)
    C.reset_axis_configs()
    ref = torch.empty_like(x)
    C.do_transform(x, ref, kinds)
    y = torch.empty_like(x)
    back = torch.empty_like(x)
    sl = (slice(0, N), slice(0, N), slice(0, N))
    scale = (N / 2.0) ** 3
    tol = AGREE_TOL[prec] if prec == 64 else 1e-5

    def run():
        C.do_transform(x, y, kinds)

    def check():
        run()
        C.do_transform(y, back, inv)
        torch.cuda.synchronize()
        rt = _rel(back[sl] / scale, x[sl])
        if not rt < tol:
            return False, f"round trip error {rt:.2e}"
        d = _rel(y[sl], ref[sl])
        if not d < tol:
            return False, f"differs from the built-in configuration by {d:.2e}"
        return True, ""

    table = []
    for role, names in [(C.ROLE_ROWS, ("x",)), (C.ROLE_COLS, ("y",)), (C.ROLE_ZCOLS, ("z",))]:
        shapes = compiled_shapes(N, prec, role) if compiled_only else None
        best, res = tune_role(N, prec, role, run, check, names, log=log, shapes=shapes, reps=reps)
        log(f"role {C.ROLE_NAMES[role]} N={N} prec={prec}: best ept={best[1]} tpl={best[2]} lines={best[3]} "
            f"{best[0]:.3f} ms ({len(res)} candidates)")
        table.append((role, N, prec, best[1], best[2], best[3], best[0], res))
    return table


def tune_batch(N, prec=64, count=32, reps=3, log=print, compiled_only=False):
    """Tune the roles of one size class of the batched solve: a batch of
    `count` N^3 problems (codes and lambda drawn like the batched workload),
    timed per launch through direct (non-graph) batched solves."""
    import numpy as np
    import torch
    from cuhelmholtz_amd import checks
    from cuhelmholtz_amd import workloads as W
    dev = torch.device("cuda:0")
    descs = W.batched_problems(count=count, seed=7, sizes=(N,))
    probs, f0, host_bcs = [], [], []
    for d in descs:
        f, bcs = W.problem_arrays(d)
        ft = torch.from_numpy(f).to(dev)
        f0.append(ft.clone())
        host_bcs.append(bcs)
        probs.append((ft, d["n"], d["bc"], d["lam"], [None if b is None else torch.from_numpy(b).to(dev) for b in bcs]))
    refine = 0 if prec == 64 else 1
    res_tol = RES_TOL[64] if prec == 64 else 1e-10   # FP32 + 1 refinement (near-resonant problems included)
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
    agree = 1e-12 if prec == 64 else 1e-9

    def check():
        solve_into().close()
        for k, (p, r, g) in enumerate(zip(probs, ref, f0)):
            res = checks.relative_residual(p[0], g, p[1], p[2], p[3], host_bcs[k])
            if not res < res_tol:
                return False, f"problem {k}: relative residual {res:.2e}"
            d = _rel(p[0], r)
            if not d < agree:
                return False, f"problem {k}: differs from the built-in configuration by {d:.2e}"
        return True, ""

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
        shapes = compiled_shapes(N, prec, role) if compiled_only else candidates(N)
        for ept, tpl, lines in shapes:
            try:
                C.set_axis_config(N, prec, role, ept, tpl, lines)
            except C.CHZError:
                continue
            ok, why = check()
            if not ok:
                log(f"  reject ept={ept} tpl={tpl} lines={lines}: {why}")
                continue
            results.append((timed(names), ept, tpl, lines))
        if not results:
            raise C.CHZError(f"no candidate of batch role {C.ROLE_NAMES[role]} N={N} passed its checks")
        results.sort()
        best = results[0]
        C.set_axis_config(N, prec, role, best[1], best[2], best[3])
        log(f"batch role {C.ROLE_NAMES[role]} N={N} prec={prec}: best ept={best[1]} tpl={best[2]} lines={best[3]} "
            f"{best[0]:.3f} ms ({len(results)} candidates; worst {results[-1][0]:.3f} ms)")
        table.append((role, N, prec, best[1], best[2], best[3], best[0], results))
    return table


def write_table(path, rows, header):
    """merge rows into a tuning table file (later rows win per (role, N, prec);
    comment lines other than the header are kept)"""
    entries, notes = {}, []
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                v = line.split("#", 1)[0].split()
                if len(v) == 6:
                    entries[tuple(int(x) for x in v[:3])] = line.rstrip("\n")
                elif line.startswith("#") and line not in header:
                    notes.append(line.rstrip("\n"))
    for r in rows:
        role, N, prec, ept, tpl, lines, ms = r[:7]
        entries[(role, N, prec)] = f"{role} {N} {prec} {ept} {tpl} {lines}   # {C.ROLE_NAMES[role]} {ms:.3f} ms (autotune)"
    with open(path, "w") as f:
        f.write(header)
        for n_ in notes:
            f.write(n_ + "\n")
        for k in sorted(entries):
            f.write(entries[k] + "\n")


def emit_alt(path, rows, k, chosen_paths):
    """Append the k fastest candidates of every tuned role that are in no
    table yet to the alternatives table (compiled by the next build)."""
    have = set()
    for p in chosen_paths:
        have.update(G.read_table(p))
    new = []
    for r in rows:
        role, N, prec = r[:3]
        for ms, ept, tpl, lines in r[7][:k]:
            row = (role, N, prec, ept, tpl, lines)
            if row not in have:
                have.add(row)
                new.append(f"{role} {N} {prec} {ept} {tpl} {lines}   # {C.ROLE_NAMES[role]} {ms:.3f} ms on the generic kernel")
    if new:
        with open(path, "a") as f:
            f.write("\n".join(new) + "\n")
    return new


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--solve", type=int, nargs=3)
    ap.add_argument("--bc", type=int, nargs=3, default=(3, 4, 0),
                    help="boundary codes of the tuned solve (a non-periodic z code tunes role 5)")
    ap.add_argument("--roles", type=int, nargs="*", help="only these roles of the tuned solve")
    ap.add_argument("--transform", type=int)
    ap.add_argument("--batch", type=int, nargs="*", help="size classes of the batched solve to tune")
    ap.add_argument("--prec", type=int, default=64)
    ap.add_argument("--arch", default=os.environ.get("CHZ_OFFLOAD_ARCH", "gfx950"))
    ap.add_argument("--out", default=None, help="tuning table to merge the winners into (default tuning/<arch>.txt)")
    ap.add_argument("--emit-alt", type=int, default=0, metavar="K",
                    help="append the K fastest candidates per role to tuning/<arch>.alt.txt instead of writing winners")
    ap.add_argument("--compiled-only", action="store_true", help="time only shapes with compiled fixed-shape kernels")
    args = ap.parse_args()
    chosen, alt = G.table_paths(args.arch)
    out = args.out or chosen
    rows = []
    if args.solve:
        rows += tune_solve(tuple(args.solve), args.prec, bc=tuple(args.bc), compiled_only=args.compiled_only,
                           roles=args.roles)
    if args.transform:
        rows += tune_transform(args.transform, args.prec, compiled_only=args.compiled_only)
    for N in args.batch or []:
        rows += tune_batch(N, args.prec, compiled_only=args.compiled_only)
    for r in rows:
        print(f"{C.ROLE_NAMES[r[0]]} N={r[1]} prec={r[2]}: top 5 " +
              "; ".join(f"{ms:.3f}ms e{e} t{t} l{l}" for ms, e, t, l in r[7][:5]))
    if args.emit_alt:
        new = emit_alt(alt, rows, args.emit_alt, (chosen, alt))
        print(f"{len(new)} shapes appended to {alt}; rebuild, then tune again with --compiled-only")
        return 0
    os.makedirs(os.path.dirname(out), exist_ok=True)
    header = ("# launch configurations chosen by cuhelmholtz_amd.autotune on MI355X (gfx950)\n"
              "# role N prec ept tpl lines   (role: 0 rows, 1 cols, 2 solve xfwd, 3 solve zmid (periodic z),\n"
              "# 4 z cols, 5 solve zmid (z codes 1-4))\n")
    write_table(out, rows, header)
    return 0


if __name__ == "__main__":
    sys.exit(main())
