"""CPU checks of the C-ABI library: it loads and exports every declared symbol
(no compute calls: this container has no GPU)."""
import ctypes
import os
import re

import pytest

import cuhelmholtz_amd as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_loads_and_exports_header_symbols():
    L = C.lib()
    names = C.header_symbols()
    assert "chz_do_transform" in names and "chz_last_error" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_cpp_dropin_symbols_exported():
    """Mangled C++ names of the reference API (funcinterface.h, cuda_helmholtz_solver.h)."""
    L = C.lib()
    for sym in ["_Z12do_transformPdS_iii14transform_kindS0_S0_",
                "_Z20do_transform_nocubicPdS_iii14transform_kindS0_S0_",
                "_Z18do_transform_cubicPdiii14transform_kindS0_S0_",
                "_Z16freeMemory_cubicv",
                "_Z21cuda_helmholtz_solverddiiPdS_ddiiS_S_ddiiS_S_dS_",
                # hw3crt_wrapper.h: FISHPACK's hw3crt_ (C linkage) and the C++ solver_3d wrapper
                "hw3crt_", "_Z9solver_3dddiiPdS_ddiiS_S_ddiiS_S_dS_"]:
        assert hasattr(L, sym), sym


def test_last_error_empty_and_version():
    L = C.lib()
    assert L.chz_last_error() == b""
    assert L.chz_version() >= 100


def test_headers_declare_reference_api():
    with open(os.path.join(ROOT, "include", "funcinterface.h")) as f:
        txt = f.read()
    for fn in ["do_transform", "do_transform_nocubic", "do_transform_cubic", "freeMemory_cubic"]:
        assert re.search(r"\b%s\s*\(" % fn, txt)
    with open(os.path.join(ROOT, "include", "global.h")) as f:
        g = f.read()
    order = re.findall(r"\b(DCT_1|DCT_2|DCT_3|DST_1|DST_2|DST_3|DFT_R2C|DFT_C2R)\b", g)
    assert order[:8] == ["DCT_1", "DCT_2", "DCT_3", "DST_1", "DST_2", "DST_3", "DFT_R2C", "DFT_C2R"]


def smooth235(m):
    for p in (2, 3, 5):
        while m % p == 0:
            m //= p
    return m == 1


def test_axis_planner_covers_235_smooth_sizes():
    """Host planner (no GPU needed): every even N with N/2 2,3,5-smooth gets a
    configuration whose radices multiply to N/2 and fit the per-thread registers."""
    for N in range(2, 2050, 2):
        M = N // 2
        for prec in (64, 32):
            for rows in (True, False):
                if not smooth235(M):
                    with pytest.raises(C.CHZError):
                        C.axis_config(N, prec, rows)
                    continue
                cfg = C.axis_config(N, prec, rows)
                prod = 1
                for R in cfg["radices"]:
                    prod *= R
                    nb = -(-(M // R) // cfg["tpl"])
                    assert nb <= cfg["ept"] // R, (N, cfg)
                assert prod == M, (N, cfg)
                assert M <= cfg["tpl"] * cfg["ept"] and cfg["tpl"] * cfg["lines"] <= 512, (N, cfg)
