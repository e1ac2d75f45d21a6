"""CPU checks of the kernel-instantiation generator and the tuning tables: the
Python radix planner agrees with the library's (chz_radix_plan, host only),
the compiled fixed-shape kernels (chz_compiled_kernels, host only) are exactly
the generator's set for the committed tables, and the tuner's table writers
keep what they should."""
import os

import pytest

import cuhelmholtz_amd as C
from cuhelmholtz_amd import autotune as A
from cuhelmholtz_amd import kernelgen as G


def test_radix_plan_matches_library():
    # every M up to 600 (powers of 2, 3, 5, large primes, mixed factors) and a
    # spread of thread shapes
    for M in range(1, 601):
        for ept in G.EPTS:
            for tpl in (1, 2, 3, 4, 5, 7, 8, 12, 16, 24, 32, 33, 64, 128, 256):
                want = C.radix_plan(M, ept, tpl)
                got = G.radix_plan(M, ept, tpl)
                assert got == want, (M, ept, tpl, got, want)


def test_radix_plan_products():
    # the stages multiply to M (byte 1 = the whole remaining factor)
    for M in (96, 128, 192, 256, 250, 343, 509, 1018):
        # the fewest threads per line with a plan (the library's built-in search)
        tpl = -(-M // 8)
        while G.radix_plan(M, 8, tpl) is None:
            tpl += 1
        p = G.radix_plan(M, 8, tpl)
        prod, rem = 1, M
        while p:
            r = p & 255
            r = rem if r == 1 else r
            prod *= r
            rem //= r
            p >>= 8
        assert prod == M


def test_compiled_kernels_are_the_generated_set():
    got = sorted(C.compiled_kernels())
    want = G.kernel_set()
    assert got == want


def test_every_chosen_row_is_compiled():
    chosen, _ = G.table_paths()
    have = {(k[1], k[2], k[3], k[4], k[5], k[6]) for k in C.compiled_kernels() if k[0] in (G.F_LINE, G.F_FIXED)}
    for row in G.read_table(chosen):
        role, N, prec, ept, tpl, lines = row
        for op in G.role_ops(role, prec):
            assert (prec, op, N, ept, tpl, lines) in have, row


def test_benchmark_shapes_are_chosen():
    # the 512^3 solve roles (FP64 and FP32) have rows
    chosen, _ = G.table_paths()
    keys = {(r[0], r[1], r[2]) for r in G.read_table(chosen)}
    for prec in (64, 32):
        for role in (G.ROLE_ROWS, G.ROLE_COLS, G.ROLE_XFWD, G.ROLE_ZMID):
            assert (role, 512, prec) in keys


def test_table_rows_rejected(tmp_path):
    p = tmp_path / "t.txt"
    p.write_text("0 512 64 8 32\n")
    with pytest.raises(ValueError):
        G.read_table(str(p))
    with pytest.raises(ValueError):
        G.check_row((0, 512, 64, 8, 8, 8))      # 256 points cannot fit 8 x 8
    with pytest.raises(ValueError):
        G.check_row((0, 512, 64, 8, 32, 32))    # 1024 threads
    with pytest.raises(ValueError):
        G.check_row((7, 64, 64, 8, 4, 16))      # no such role
    assert G.check_row((3, 512, 64, 8, 32, 16)) == 0x40808


def test_generated_sources(tmp_path):
    out = G.generate(str(tmp_path))
    names = sorted(os.path.basename(p) for p in out)
    assert "chz_gen_index.cpp" in names
    text = "".join(open(p).read() for p in out)
    # one entry per kernel, each kernel instantiated once
    assert text.count("(const void *)") == len(G.kernel_set())
    # a second run leaves unchanged files alone
    mt = {p: os.path.getmtime(p) for p in out}
    G.generate(str(tmp_path))
    assert all(os.path.getmtime(p) == mt[p] for p in out)


def test_emit_alt_and_write_table(tmp_path):
    chosen = tmp_path / "c.txt"
    alt = tmp_path / "a.txt"
    chosen.write_text("# head\n3 64 64 8 4 32   # zmid\n")
    alt.write_text("3 64 64 8 4 16\n")
    res = [(0.1, 8, 4, 32), (0.2, 8, 4, 16), (0.3, 8, 8, 8), (0.4, 4, 8, 16)]
    rows = [(3, 64, 64, 8, 4, 32, 0.1, res)]
    new = A.emit_alt(str(alt), rows, 3, (str(chosen), str(alt)))
    assert len(new) == 1 and new[0].startswith("3 64 64 8 8 8")
    assert G.read_table(str(alt)) == [(3, 64, 64, 8, 4, 16), (3, 64, 64, 8, 8, 8)]
    A.write_table(str(chosen), [(3, 64, 64, 8, 8, 8, 0.3, res), (1, 64, 64, 8, 4, 16, 0.2, res)], "# head\n")
    assert G.read_table(str(chosen)) == [(1, 64, 64, 8, 4, 16), (3, 64, 64, 8, 8, 8)]


def test_zmid_role_specialisation(tmp_path):
    # a role-5 row (fused z pass, non-periodic z) compiles no per-code kernels;
    # the role-3 row of the same size is then specialised for code 0 only,
    # without a role-5 row for every code
    chosen = tmp_path / "c.txt"
    alt = tmp_path / "a.txt"
    alt.write_text("")
    chosen.write_text("3 512 64 16 16 16\n5 512 64 8 32 16\n3 512 32 16 16 32\n")
    ks = G.kernel_set([str(chosen), str(alt)])
    spec = {(k[1], k[4], k[5], k[6]): set() for k in ks}
    for k in ks:
        if k[8] >= 0:
            spec[(k[1], k[4], k[5], k[6])].add(k[8])
    assert spec[(64, 16, 16, 16)] == {0}
    assert spec[(64, 8, 32, 16)] == set()
    assert spec[(32, 16, 16, 32)] == {0, 1, 2, 3, 4}
    assert all(k[2] == G.B_ZMID for k in ks)


def test_role5_rows_accepted():
    assert G.check_row((5, 512, 64, 8, 32, 16)) == 0x40808
    assert G.role_ops(G.ROLE_ZMIDN, 32) == [G.B_ZMID]
