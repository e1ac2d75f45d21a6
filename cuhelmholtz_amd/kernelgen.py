"""Kernel instantiation tables generated from the tuning tables.

The fixed-shape pass kernels have every launch parameter (cells per line N,
points per thread, threads per line, lines per tile, radix plan) as a
compile-time constant, so each shape is a separate kernel.  Which shapes are
compiled is decided here, from the autotuner's output, not by hand:

  tuning/<arch>.txt      the launch configuration the autotuner chose per
                         (role, N, precision); the library loads it at start-up
  tuning/<arch>.alt.txt  further shapes the autotuner measures against the
                         chosen ones (written by `autotune --emit-alt`)

Every row "role N prec ept tpl lines" of either file is compiled for each pass
the role runs (rows: the x inverse pass and the transform's x pass, plus the
mixed solve's inverse-x-and-update pass in FP32; cols: y passes and the
transform's y pass; xfwd: solve pass 1, plus the FP64 residual pass in FP32;
zmid: the fused z pass; zcols: the transform's z pass):

  * an even thread count per line: the lean tile I/O kernel k_line, and for
    N <= 128 (the batched size classes) its batched form
    k_line_batch, so a batched problem and its single solve run one tile body
    (FP32 batches end with the update pass, so their rows pass is not built);
  * an odd thread count: k_fixed, and for N <= 128 k_batch_fixed.

Rows of the chosen table with N in SPECIALIZE_N (512, the large single
solves and transforms) are also compiled once per kind the pass can run
(single-problem lean-I/O form): every kind for the rows and columns passes,
the forward kinds of the five boundary codes for solve pass 1 and the
residual pass, their inverse kinds for the update pass, and the five codes'
(forward, inverse) pairs for the fused z pass.  Such a kernel holds only that
kind's code (one with every kind inlined is several hundred KB, and its speed
changed by half with where its hot path landed in the instruction cache); the
library picks it for a pass of that kind, else the kernel for every kind.

Shapes in neither file run on the generic kernels (run-time radix plan).
build() calls generate() before compiling; a CPU test checks that the
library's compiled shapes (chz_compiled_kernels) are exactly kernel_set() of
the committed tables, and that radix_plan() agrees with the library's planner
(chz_radix_plan).
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING = os.path.join(HERE, "tuning")

B_ROWS, B_COLS, B_XFWD, B_ZMID, B_XRES, B_XUPD = range(6)
OP_NAMES = ["B_ROWS", "B_COLS", "B_XFWD", "B_ZMID", "B_XRES", "B_XUPD"]
# roles of the tuning tables (ROLE_ZMIDN: the fused z pass for non-periodic z)
ROLE_ROWS, ROLE_COLS, ROLE_XFWD, ROLE_ZMID, ROLE_ZCOLS, ROLE_ZMIDN = range(6)
# forms of a compiled shape (FixedEntry::form)
F_LINE, F_LINE_BATCH, F_FIXED, F_FIXED_BATCH = range(4)
FORM_NAMES = ["line", "line_batch", "fixed", "fixed_batch"]
EPTS = (16, 12, 10, 8, 6, 5, 4)
MAX_THREADS = 512
BATCH_MAX_N = 128
LDS_BYTES = 163840
RADICES = (16, 12, 10, 8, 6, 5, 4, 3, 2)
KERNELS_PER_FILE = 6
SPECIALIZE_N = (512,)
K_DCT_1, K_DCT_2, K_DCT_3, K_DST_1, K_DST_2, K_DST_3, K_R2C, K_C2R = range(8)
FWD_KINDS = (K_R2C, K_DST_1, K_DST_3, K_DCT_1, K_DCT_3)   # boundary codes 0..4
INV_KINDS = (K_C2R, K_DST_1, K_DST_2, K_DCT_1, K_DCT_2)


def kind_sels(op):
    """Kind selectors a pass op can run with (the fused z pass: boundary codes)."""
    if op in (B_ROWS, B_COLS):
        return tuple(range(8))
    if op in (B_XFWD, B_XRES):
        return FWD_KINDS
    if op == B_XUPD:
        return INV_KINDS
    return tuple(range(5))


def radix_plan(M, ept, tpl):
    """Radix plan of an M-point FFT with tpl threads per line and at most ept
    points per thread, packed one byte per stage (first stage in the low byte),
    or None if the shape cannot run.  Restates radix_plan() of
    csrc/chz_context.cpp: the largest specialised radix R <= ept dividing the
    remaining factor whose butterflies fit (ceil(M/(R tpl)) <= ept/R); a
    factor without one (primes from 7 up) becomes a direct stage (byte R, or
    byte 1 for a remaining factor above 255)."""
    if M < 1 or M > ept * tpl:
        return None
    plan, rem, st = 0, M, 0
    while rem > 1:
        pick = 0
        for R in RADICES:
            if R > ept or rem % R:
                continue
            if -(-(M // R) // tpl) > ept // R:
                continue
            pick = R
            break
        if not pick:
            q = 2
            while q * q <= rem and rem % q:
                q += 1
            if rem % q:
                q = rem
            if q < 7:
                return None
            pick = q if q <= 255 else rem
        if st >= 8:
            return None
        plan |= (pick if pick <= 255 else 1) << (8 * st)
        st += 1
        rem //= pick
    return plan


def axis_pass_lds(prec, M, L, tpl):
    """LDS bytes of a pass tile (axis_pass_lds in csrc/chz_launch.cpp)."""
    esz = 8 if prec == 64 else 4
    ls = L + (0 if esz == 8 else 1)
    return ((M + 1) * ls + M) * 2 * esz + tpl * L * esz


def role_ops(role, prec):
    """The passes (BatchOp) that run with a role's launch configuration."""
    ops = {ROLE_ROWS: [B_ROWS], ROLE_COLS: [B_COLS], ROLE_XFWD: [B_XFWD], ROLE_ZMID: [B_ZMID],
           ROLE_ZCOLS: [B_COLS], ROLE_ZMIDN: [B_ZMID]}[role]
    if prec == 32 and role == ROLE_ROWS:
        ops = ops + [B_XUPD]
    if prec == 32 and role == ROLE_XFWD:
        ops = ops + [B_XRES]
    return ops


def read_table(path):
    """Rows (role, N, prec, ept, tpl, lines) of a tuning table ('#' comments)."""
    rows = []
    if not os.path.exists(path):
        return rows
    with open(path) as f:
        for ln, line in enumerate(f, 1):
            body = line.split("#", 1)[0].split()
            if not body:
                continue
            if len(body) != 6:
                raise ValueError(f"{path}:{ln}: expected 'role N prec ept tpl lines'")
            rows.append(tuple(int(v) for v in body))
    return rows


def check_row(row, where=""):
    """The plan of a table row; ValueError if the shape cannot run (the
    checks of Context::set_config)."""
    role, N, prec, ept, tpl, lines = row
    M = N // 2
    if role not in range(6) or prec not in (32, 64) or ept not in EPTS or N < 2 or N % 2:
        raise ValueError(f"{where}invalid tuning row {row}")
    if tpl < 1 or lines < 1 or tpl * lines > MAX_THREADS:
        raise ValueError(f"{where}tuning row {row}: threads per tile out of range")
    plan = radix_plan(M, ept, tpl)
    if plan is None:
        raise ValueError(f"{where}tuning row {row}: no radix plan")
    if axis_pass_lds(prec, M, lines, tpl) > LDS_BYTES:
        raise ValueError(f"{where}tuning row {row}: tile does not fit in LDS")
    return plan


def table_paths(arch="gfx950"):
    return [os.path.join(TUNING, f"{arch}.txt"), os.path.join(TUNING, f"{arch}.alt.txt")]


def kernel_set(paths=None, arch="gfx950"):
    """Sorted, de-duplicated kernels to compile: (form, prec, op, N, ept, tpl,
    lines, plan, ks), ks = -1 for every kind, else the kind selector."""
    out = set()
    paths = paths or table_paths(arch)
    # sizes whose non-periodic fused z pass has a chosen shape of its own
    zmidn = {(r[1], r[2]) for r in read_table(paths[0]) if r[0] == ROLE_ZMIDN}
    for pi, path in enumerate(paths):
        for row in read_table(path):
            role, N, prec, ept, tpl, lines = row
            plan = check_row(row, f"{os.path.basename(path)}: ")
            single, batched = (F_LINE, F_LINE_BATCH) if tpl % 2 == 0 else (F_FIXED, F_FIXED_BATCH)
            for op in role_ops(role, prec):
                out.add((single, prec, op, N, ept, tpl, lines, plan, -1))
                # batched solves end their FP32 passes with the update pass, not a rows pass
                if N <= BATCH_MAX_N and not (prec == 32 and op == B_ROWS):
                    out.add((batched, prec, op, N, ept, tpl, lines, plan, -1))
                # kind-specialised forms of the chosen large shapes (first table
                # only); the non-periodic fused z pass runs the kernel for every
                # kind (its per-code FP32 kernels spilled), so where it has a row
                # the periodic role's shape is specialised for code 0 only
                if pi == 0 and N in SPECIALIZE_N and single == F_LINE and role != ROLE_ZMIDN:
                    sels = (0,) if role == ROLE_ZMID and (N, prec) in zmidn else kind_sels(op)
                    for ks in sels:
                        out.add((single, prec, op, N, ept, tpl, lines, plan, ks))
    return sorted(out)


def _entry(k):
    form, prec, op, N, ept, tpl, lines, plan, ks = k
    T = "double" if prec == 64 else "float"
    opn = OP_NAMES[op]
    pl = f"0x{plan:x}ull"
    if form == F_LINE:
        fn = f"k_line<{T}, {opn}, {N}, {tpl}, {lines}, {pl}, {ept}" + (f", {ks}>" if ks >= 0 else ">")
    elif form == F_LINE_BATCH:
        fn = f"k_line_batch<{T}, {opn}, {N}, {tpl}, {lines}, {pl}, {ept}" + (f", {ks}>" if ks >= 0 else ">")
    elif form == F_FIXED:
        fn = f"k_fixed<{T}, {ept}, {opn}, {N}, {tpl}, {lines}, {pl}>"
    else:
        fn = f"k_batch_fixed<{T}, {ept}, {opn}, {N}, {tpl}, {lines}, {pl}>"
    if form == F_LINE:
        extra = f"line_lds_extra<{T}, {N}, {opn}, {tpl}, {lines}>()"
    elif form == F_LINE_BATCH:
        extra = f"line_lds_extra<{T}, {N}, {opn}, {tpl}, {lines}, true>()"
    else:
        extra = "0u"
    return (f"    {{{prec}, {opn}, {N}, {ept}, {tpl}, {lines}, {pl}, (const void *){fn},\n"
            f"     {extra}, F_{FORM_NAMES[form].upper()}, {ks}}},")


def sources(paths=None, arch="gfx950"):
    """{file name: text} of the generated translation units."""
    ks = kernel_set(paths, arch)
    # interleave forms and sizes over the files so their compile times even out
    nfiles = max(1, -(-len(ks) // KERNELS_PER_FILE))
    chunks = [ks[i::nfiles] for i in range(nfiles)]
    head = ("// Generated by cuhelmholtz_amd/kernelgen.py from the tuning tables\n"
            "// (" + ", ".join(os.path.relpath(p, HERE) for p in (paths or table_paths(arch))) + "); do not edit.\n")
    out = {}
    for i, ch in enumerate(chunks):
        body = "\n".join(_entry(k) for k in ch)
        out[f"chz_gen_{i:02d}.hip"] = (
            head + '#include "chz_line.h"\n\nnamespace chz {\n\nstatic const FixedEntry kTab[] = {\n' + body +
            "\n};\n\nconst FixedEntry *gen_table_%02d(int *n)\n{\n    *n = (int)(sizeof(kTab) / sizeof(kTab[0]));\n"
            "    return kTab;\n}\n\n}  // namespace chz\n" % i)
    decl = "".join(f"const FixedEntry *gen_table_{i:02d}(int *n);\n" for i in range(len(chunks)))
    cases = "".join(f"        case {i}: return gen_table_{i:02d}(n);\n" for i in range(len(chunks)))
    out["chz_gen_index.cpp"] = (
        head + '#include "chz_kernels.h"\n\nnamespace chz {\n\n' + decl +
        f"\nint gen_table_count() {{ return {len(chunks)}; }}\n\n"
        "const FixedEntry *gen_table(int q, int *n)\n{\n    switch (q) {\n" + cases +
        "    }\n    *n = 0;\n    return nullptr;\n}\n\n}  // namespace chz\n")
    return out


def generate(out_dir, paths=None, arch="gfx950"):
    """Write the generated sources into out_dir (files whose text is unchanged
    are left alone, so an incremental build does not recompile them; stale
    generated files are removed).  Returns their paths."""
    os.makedirs(out_dir, exist_ok=True)
    srcs = sources(paths, arch)
    for f in os.listdir(out_dir):
        if f.startswith("chz_gen_") and f not in srcs:
            os.remove(os.path.join(out_dir, f))
    written = []
    for name, text in sorted(srcs.items()):
        p = os.path.join(out_dir, name)
        old = None
        if os.path.exists(p):
            with open(p) as f:
                old = f.read()
        if old != text:
            with open(p, "w") as f:
                f.write(text)
        written.append(p)
    return written


if __name__ == "__main__":
    for k in kernel_set():
        print(FORM_NAMES[k[0]], k[1:7], hex(k[7]), k[8])
