"""GPU parity of the FP64 Helmholtz solve (HIP, through the C ABI).

Bar: 1e-12 relative L-infinity against the oracle's restatement of
cuda_helmholtz_solver, against sparse direct solves (golden fixtures) and
against the hw3crt-contract CPU solver; O(h^2) against the analytic
manufactured solution; at the full benchmark sizes (512^3, 512x384x256) the
normalised FP64 7-point residual |f - A u| / (|A| |u| + |f|) < 1e-13.
"""
import itertools
import os

import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not torch.cuda.is_available(), reason="no GPU")]

import cuhelmholtz_amd as C  # noqa: E402
from cuhelmholtz_amd import workloads as W  # noqa: E402
from cuhelmholtz_amd import checks as H  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-12


def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def faces(rng, n, bc):
    NX, NY, NZ = n
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    return [rng.standard_normal(sizes[i]) if bc[i // 2] in ((3, 4) if i % 2 == 0 else (2, 3)) else None
            for i in range(6)]


def gpu_solve(f, n, bc, lam, bcs, dom=(0.0, 1.0) * 3):
    ft = torch.from_numpy(np.ascontiguousarray(f)).to(dev())
    C.helmholtz_solve(ft, n, bc, lam, bcs, dom)
    torch.cuda.synchronize()
    return ft.cpu().numpy()


def test_golden_sparse_direct():
    g = np.load(os.path.join(GOLD, "golden_solver.npz"))
    n = tuple(int(v) for v in g["n"])
    for t in range(int(g["count"])):
        bc = tuple(int(b) for b in g[f"bc_{t}"])
        lam = float(g[f"lam_{t}"])
        bcs = [g[f"{nm}_{t}"] if f"{nm}_{t}" in g else None for nm in ["bcl", "bcr", "bcb", "bct", "bcf", "bce"]]
        got = gpu_solve(g[f"f_{t}"], n, bc, lam, bcs)
        want = g[f"u_{t}"] if f"u_{t}" in g else O.helmholtz(g[f"f_{t}"], n, bc, lam, bcs)
        assert rel(got, want) < TOL, (bc, lam)


@pytest.mark.parametrize("n,dom", [((6, 8, 10), (0.0, 1.0) * 3), ((12, 4, 6), (0.0, 2.0, -1.0, 0.5, 0.25, 1.0)),
                                   ((16, 24, 8), (0.0, 1.0) * 3), ((14, 22, 26), (0.0, 1.0) * 3),
                                   ((34, 28, 42), (0.0, 1.5, 0.0, 1.0, -1.0, 1.0))])
def test_all_bc_vs_oracle(n, dom):
    """Every code triple against the oracle, including sizes whose N/2 has
    prime factors 7, 11, 13, 17 (direct FFT stages)."""
    rng = np.random.default_rng(1)
    for bc in itertools.product(range(5), repeat=3):
        f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
        bcs = faces(rng, n, bc)
        got = gpu_solve(f, n, bc, -2.3, bcs, dom)
        want = O.helmholtz(f, n, bc, -2.3, bcs, dom)
        assert rel(got, want) < TOL, bc


@pytest.mark.parametrize("n,dom", [((6, 8, 10), (0.0, 1.0) * 3), ((12, 4, 6), (0.0, 2.0, -1.0, 0.5, 0.25, 1.0))])
def test_19point_all_bc_vs_oracle(n, dom):
    """The 19-point operator (chz_helmholtz_solve19) for every code triple, with
    boundary data folded, against the oracle (pinned by the physical-space stencil)."""
    rng = np.random.default_rng(19)
    for bc in itertools.product(range(5), repeat=3):
        for lam in (-2.3, 0.0):
            f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
            bcs = faces(rng, n, bc)
            ft = torch.from_numpy(f.copy()).to(dev())
            C.helmholtz_solve(ft, n, bc, lam, bcs, dom, stencil=19)
            torch.cuda.synchronize()
            want = O.helmholtz(f, n, bc, lam, bcs, dom, stencil=19)
            assert rel(ft.cpu().numpy(), want) < TOL, (bc, lam)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_long_axis_every_bc_vs_oracle(axis):
    """512 cells along one axis (the fixed-shape pass kernels of that axis's
    role: x rows, y columns, fused z pass) with every code on it, partial
    tiles on the short axes, against the oracle."""
    rng = np.random.default_rng(40 + axis)
    n = [6, 8, 10]
    n[axis] = 512
    n = tuple(n)
    for code in range(5):
        bc = [int(v) for v in rng.integers(0, 5, 3)]
        bc[axis] = code
        bc = tuple(bc)
        for lam in (-2.3, 0.0):
            f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
            bcs = faces(rng, n, bc)
            got = gpu_solve(f, n, bc, lam, bcs)
            want = O.helmholtz(f, n, bc, lam, bcs)
            assert rel(got, want) < TOL, (n, bc, lam)
            # the 19-point operator through the same pass kernels
            ft = torch.from_numpy(f.copy()).to(dev())
            C.helmholtz_solve(ft, n, bc, lam, bcs, stencil=19)
            torch.cuda.synchronize()
            assert rel(ft.cpu().numpy(), O.helmholtz(f, n, bc, lam, bcs, stencil=19)) < TOL, (n, bc, lam, 19)


ZMID_512_CASES = [((18, 6, 512), (3, 1, 0), -2.3), ((32, 8, 512), (0, 0, 0), 0.0), ((32, 8, 512), (0, 4, 0), 0.5),
                  ((18, 6, 512), (2, 3, 4), -1.0), ((10, 6, 512), (4, 2, 1), -7.0), ((32, 6, 512), (1, 0, 3), 0.0),
                  ((16, 6, 512), (3, 4, 2), -3.1)]


def test_zmid_512_every_compiled_shape_vs_oracle():
    """Every compiled launch shape of the 512-cell fused z pass (FP64), the
    wave-resident periodic form (16 x 16 points per line) among them, as the
    configuration of both fused-z roles (periodic z and z codes 1-4), on z
    lines of every code, full and partial 16-column tiles (NX = 18 with a
    19-column DCT-I window, NX = 32 periodic), the singular periodic /
    Neumann case and the 19-point operator, against the oracle."""
    from cuhelmholtz_amd import autotune as A
    shapes = A.compiled_shapes(512, 64, C.ROLE_ZMID)
    assert (16, 16, 16) in shapes and (8, 32, 16) in shapes
    rng = np.random.default_rng(77)
    try:
        for ept, tpl, lines in shapes:
            for role in (C.ROLE_ZMID, C.ROLE_ZMIDN):
                C.set_axis_config(512, 64, role, ept, tpl, lines)
            for n, bc, lam in ZMID_512_CASES:
                f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
                bcs = faces(rng, n, bc)
                for stencil in (7, 19):
                    ft = torch.from_numpy(f.copy()).to(dev())
                    C.helmholtz_solve(ft, n, bc, lam, bcs, stencil=stencil)
                    torch.cuda.synchronize()
                    want = O.helmholtz(f, n, bc, lam, bcs, stencil=stencil)
                    assert rel(ft.cpu().numpy(), want) < TOL, ((ept, tpl, lines), n, bc, lam, stencil)
    finally:
        C.reset_axis_configs()
        C.load_tuning()


def test_zmid_512_fp32_every_compiled_shape_vs_oracle():
    """The same for the FP32 fused z pass of the mixed solve (two FP64
    refinements: some of the cases have spectral condition numbers above 1e5,
    where one leaves a few 1e-10; bar 1e-10 against the oracle)."""
    from cuhelmholtz_amd import autotune as A
    shapes = A.compiled_shapes(512, 32, C.ROLE_ZMID)
    assert (8, 32, 16) in shapes
    rng = np.random.default_rng(78)
    try:
        for ept, tpl, lines in shapes:
            for role in (C.ROLE_ZMID, C.ROLE_ZMIDN):
                C.set_axis_config(512, 32, role, ept, tpl, lines)
            for n, bc, lam in ZMID_512_CASES:
                f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
                bcs = faces(rng, n, bc)
                ft = torch.from_numpy(f.copy()).to(dev())
                C.helmholtz_solve_mixed(ft, n, bc, lam, bcs, refine=2)
                torch.cuda.synchronize()
                want = O.helmholtz(f, n, bc, lam, bcs)
                assert rel(ft.cpu().numpy(), want) < 1e-10, ((ept, tpl, lines), n, bc, lam)
    finally:
        C.reset_axis_configs()
        C.load_tuning()


def test_zmid_roles():
    """The non-periodic fused z pass (role 5) uses the periodic one's
    configuration (role 3) until it has one of its own, also after role 3
    changes; a role-5 configuration is then its own, and a solve with
    non-periodic z runs with it (parity against the oracle)."""
    rng = np.random.default_rng(79)
    try:
        C.set_axis_config(64, 64, C.ROLE_ZMID, 8, 4, 16)
        assert C.axis_config(64, 64, role=C.ROLE_ZMIDN) == C.axis_config(64, 64, role=C.ROLE_ZMID)
        C.set_axis_config(64, 64, C.ROLE_ZMID, 8, 4, 32)
        assert C.axis_config(64, 64, role=C.ROLE_ZMIDN)["lines"] == 32
        C.set_axis_config(64, 64, C.ROLE_ZMIDN, 4, 8, 8)
        assert C.axis_config(64, 64, role=C.ROLE_ZMIDN)["ept"] == 4
        assert C.axis_config(64, 64, role=C.ROLE_ZMID)["lines"] == 32
        for bc in ((3, 4, 0), (3, 4, 2)):
            n = (24, 16, 64)
            f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
            bcs = faces(rng, n, bc)
            assert rel(gpu_solve(f, n, bc, -2.0, bcs), O.helmholtz(f, n, bc, -2.0, bcs)) < TOL, bc
    finally:
        C.reset_axis_configs()
        C.load_tuning()


@pytest.mark.parametrize("bc", [(0, 0, 0), (3, 3, 3), (0, 3, 0), (3, 0, 3)])
def test_singular_zero_mode(bc):
    rng = np.random.default_rng(2)
    n = (8, 12, 6)
    f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
    bcs = faces(rng, n, bc)
    got = gpu_solve(f, n, bc, 0.0, bcs)
    assert rel(got, O.helmholtz(f, n, bc, 0.0, bcs)) < TOL
    v, ierr, _ = O.solver_3d(f, n, bc, 0.0, bcs)
    assert rel(got, v) < TOL


def test_manufactured_vs_hw3crt_and_second_order():
    """The manufactured-solution workload (64^3, codes (1,2,0), lambda -1) against
    the hw3crt-contract CPU solver at 1e-12, and O(h^2) against the analytic u."""
    errs = []
    for N in (32, 64, 128):
        f, u, bcs = W.manufactured((N, N, N))
        got = gpu_solve(f, (N, N, N), (1, 2, 0), -1.0, bcs)
        if N == 64:
            v, ierr, _ = O.solver_3d(f, (N, N, N), (1, 2, 0), -1.0, bcs, threads=8)
            assert ierr == 0 and rel(got, v) < TOL
        errs.append(np.max(np.abs(got - u)))
    assert 3.6 < errs[0] / errs[1] < 4.4 and 3.6 < errs[1] / errs[2] < 4.4, errs


@pytest.mark.parametrize("n,bc", [((128, 128, 128), (3, 4, 0)), ((256, 192, 128), (2, 1, 3))])
def test_vs_hw3crt_medium(n, bc):
    rng = np.random.default_rng(3)
    f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
    bcs = faces(rng, n, bc)
    got = gpu_solve(f, n, bc, -4.0, bcs)
    v, ierr, _ = O.solver_3d(f, n, bc, -4.0, bcs, threads=16)
    assert ierr == 0 and rel(got, v) < TOL


@pytest.mark.parametrize("n", [(512, 512, 512), (512, 384, 256), (256, 256, 256), (384, 384, 384)])
@pytest.mark.parametrize("lam", [-4.0, 0.5])
def test_full_size_residual(n, lam):
    """Mixed-precision workload shape, plain FP64 path: residual invariant."""
    f, bcs = W.smooth_modes(n)
    ft = torch.from_numpy(f).to(dev())
    del f
    ff = H.fold(ft, n, (3, 4, 0), bcs)
    C.helmholtz_solve(ft, n, (3, 4, 0), lam, bcs)
    r, scale = H.residual(ft, ff, n, (3, 4, 0), lam)
    assert r / scale < 1e-13, (r, scale)
    del ft, ff
    torch.cuda.empty_cache()


def test_errors_reported():
    f = torch.zeros((9, 9, 9), dtype=torch.float64, device=dev())
    with pytest.raises(C.CHZError, match="boundary condition"):
        C.helmholtz_solve(f, (8, 8, 8), (5, 1, 1), -1.0)
    import ctypes
    L = C.lib()
    rc = L.chz_helmholtz_solve(0.0, 1.0, 8, 3, None, None, 0.0, 1.0, 8, 1, None, None, 0.0, 1.0, 8, 1, None, None,
                               -1.0, ctypes.c_void_p(f.data_ptr()), None)
    assert rc == -1 and b"Neumann" in L.chz_last_error()


def test_cpp_dropin_cuda_helmholtz_solver():
    import ctypes
    L = C.lib()
    fn = getattr(L, "_Z21cuda_helmholtz_solverddiiPdS_ddiiS_S_ddiiS_S_dS_")
    d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    fn.argtypes = [d, d, i, i, vp, vp] * 3 + [d, vp]
    fn.restype = i
    rng = np.random.default_rng(4)
    n, bc = (16, 12, 8), (4, 0, 2)
    f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
    bcs = faces(rng, n, bc)
    want = O.helmholtz(f, n, bc, -1.0, bcs)
    ft = torch.from_numpy(f).to(dev())
    bt = [None if b is None else torch.from_numpy(b).to(dev()) for b in bcs]
    p = [None if b is None else vp(b.data_ptr()) for b in bt]
    torch.cuda.synchronize()
    rc = fn(0.0, 1.0, n[0], bc[0], p[0], p[1], 0.0, 1.0, n[1], bc[1], p[2], p[3], 0.0, 1.0, n[2], bc[2], p[4], p[5],
            -1.0, vp(ft.data_ptr()))
    assert rc == 0
    assert rel(ft.cpu().numpy(), want) < TOL


# ---------------------------------------------------------------------------
# FP32 transforms + FP64 iterative refinement, against the plain FP64 solve
# ---------------------------------------------------------------------------
def mixed_solve(f, n, bc, lam, bcs, refine, dom=(0.0, 1.0) * 3):
    ft = torch.from_numpy(np.ascontiguousarray(f)).to(dev())
    C.helmholtz_solve_mixed(ft, n, bc, lam, bcs, dom, refine=refine)
    torch.cuda.synchronize()
    return ft.cpu().numpy()


@pytest.mark.parametrize("n", [(6, 8, 10), (16, 12, 8), (14, 22, 26)])
def test_mixed_all_bc_converges(n):
    """Error vs the FP64 solve: FP32 alone ~1e-6, one refinement < 1e-10, two < 1e-12."""
    rng = np.random.default_rng(11)
    for bc in itertools.product(range(5), repeat=3):
        f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
        bcs = faces(rng, n, bc)
        ref = gpu_solve(f, n, bc, -2.3, bcs)
        e0 = rel(mixed_solve(f, n, bc, -2.3, bcs, 0), ref)
        e1 = rel(mixed_solve(f, n, bc, -2.3, bcs, 1), ref)
        e2 = rel(mixed_solve(f, n, bc, -2.3, bcs, 2), ref)
        assert e0 < 1e-4 and e1 < 1e-10 and e2 < 1e-12, (bc, e0, e1, e2)
        assert e1 < e0


def test_mixed_singular_and_periodic():
    rng = np.random.default_rng(12)
    n = (8, 12, 6)
    for bc in [(0, 0, 0), (3, 3, 3), (0, 3, 0)]:
        f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
        bcs = faces(rng, n, bc)
        ref = gpu_solve(f, n, bc, 0.0, bcs)
        assert rel(mixed_solve(f, n, bc, 0.0, bcs, 2), ref) < 1e-12, bc


@pytest.mark.parametrize("n", [(512, 384, 256), (512, 512, 512), (256, 256, 256), (384, 384, 384)])
@pytest.mark.parametrize("lam", [-4.0, 0.5])
def test_mixed_full_size_1e10(n, lam):
    """Mixed-precision workload: FP32 + refinement reaches 1e-10 relative error
    against the plain FP64 solve within two refinement steps."""
    f, bcs = W.smooth_modes(n)
    ref = torch.from_numpy(f).to(dev())
    C.helmholtz_solve(ref, n, (3, 4, 0), lam, bcs)
    errs = []
    for refine in (1, 2):
        ft = torch.from_numpy(f).to(dev())
        C.helmholtz_solve_mixed(ft, n, (3, 4, 0), lam, bcs, refine=refine)
        errs.append(((ft - ref).abs().max() / ref.abs().max()).item())
        del ft
    assert min(errs) < 1e-10, errs
    del ref
    torch.cuda.empty_cache()
