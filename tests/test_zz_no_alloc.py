"""The allocation-free FP64 solve (chz_helmholtz_solve_noalloc) through the C
ABI: the same passes as chz_helmholtz_solve with the workspace and tables in a
fixed device array, so it is checked against that entry point (and the
oracle's restatement of cuda_helmholtz_solver) at 1e-12 relative
L-infinity, on two calls in a row with the second smaller than the first.
"""
import ctypes

import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not torch.cuda.is_available(), reason="no GPU")]

import cuhelmholtz_amd as C  # noqa: E402

TOL = 1e-12


def faces(rng, n, bc):
    NX, NY, NZ = n
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    return [rng.standard_normal(sizes[i]) if bc[i // 2] in ((3, 4) if i % 2 == 0 else (2, 3)) else None
            for i in range(6)]


def call(fn, u, n, bc, lam, bt, dom=(0.0, 1.0) * 3):
    args = []
    for a in range(3):
        lo, hi = bt[2 * a], bt[2 * a + 1]
        args += [float(dom[2 * a]), float(dom[2 * a + 1]), int(n[a]), int(bc[a]),
                 None if lo is None else ctypes.c_void_p(lo.data_ptr()),
                 None if hi is None else ctypes.c_void_p(hi.data_ptr())]
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return fn(*args, float(lam), ctypes.c_void_p(u.data_ptr()), s)


def test_noalloc_matches_solve_two_calls():
    L = C.lib()
    dev = torch.device("cuda:0")
    cases = [((48, 40, 64), (3, 4, 0), -2.0, (0.0, 1.0, -0.5, 0.5, 0.0, 2.0)),
             ((24, 16, 20), (0, 2, 1), 0.5, (0.0, 1.0) * 3)]
    for k, (n, bc, lam, dom) in enumerate(cases):
        rng = np.random.default_rng(70 + k)
        f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
        bcs = faces(rng, n, bc)
        bt = [None if b is None else torch.from_numpy(b).to(dev) for b in bcs]
        want = O.helmholtz(f, n, bc, lam, bcs, dom)
        ref = torch.from_numpy(f).to(dev)
        got = torch.from_numpy(f).to(dev)
        assert call(L.chz_helmholtz_solve, ref, n, bc, lam, bt, dom) == 0, L.chz_last_error()
        assert call(L.chz_helmholtz_solve_noalloc, got, n, bc, lam, bt, dom) == 0, L.chz_last_error()
        torch.cuda.synchronize()
        e_ref = float((got - ref).abs().max() / ref.abs().max())
        e_or = float(np.max(np.abs(got.cpu().numpy() - want)) / np.max(np.abs(want)))
        print(f"call {k} {n} {bc}: vs chz_helmholtz_solve {e_ref:.2e}, vs oracle {e_or:.2e}")
        assert e_ref <= TOL and e_or <= TOL, (k, e_ref, e_or)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    L = C.lib()
    n, bc = (256, 256, 256), (1, 1, 1)
    u = torch.zeros((n[2] + 1, n[1] + 1, n[0] + 1), dtype=torch.float64, device="cuda:0")
    rc = call(L.chz_helmholtz_solve_noalloc, u, n, bc, -1.0, [None] * 6)
    assert rc == -1
    assert b"CHZ_NOALLOC_SCRATCH_BYTES" in L.chz_last_error()
    assert float(u.abs().max()) == 0.0
