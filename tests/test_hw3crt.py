"""The reference's CPU entry points served by the GPU solve: hw3crt_ / solver_3d
(cudahelmholtz/hw3crt_wrapper.h:1-19, hw3crt_wrapper.cpp:5-33), and the
command-line drivers with the arguments and output of the reference's test
programs (cudahelmholtz/testgpu.cu, testgpurectangle.cu, testcpu.cpp,
testcpurectangle.cpp).

Checked against the oracle's hw3crt-contract CPU solver (oracle/chz_oracle.c):
IERROR for every invalid argument (host-side validation, no GPU needed), the
solution and PERTRB for every boundary-code triple including the singular
ones (1e-12 relative), FISHPACK's leading dimensions (LDIMF > L+1, MDIMF >
M+1), and the drivers' "max diff" against the same discrete problem solved by
the oracle.
"""
import itertools
import os
import re
import subprocess

import numpy as np
import pytest

import cuhelmholtz_amd as C
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cuhelmholtz_amd", "bin")
TOL = 1e-12


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def faces(rng, n, bc):
    NX, NY, NZ = n
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    return [rng.standard_normal(sizes[i]) if bc[i // 2] in ((3, 4) if i % 2 == 0 else (2, 3)) else None
            for i in range(6)]


# (argument changes, IERROR) as documented for FISHPACK's hw3crt
BAD = [
    (dict(dom=(1.0, 0.0, 0.0, 1.0, 0.0, 1.0)), 1),
    (dict(n=(2, 8, 8)), 2),
    (dict(bc=(5, 1, 1)), 3),
    (dict(dom=(0.0, 1.0, 1.0, 1.0, 0.0, 1.0)), 4),
    (dict(n=(8, 2, 8)), 5),
    (dict(bc=(1, -1, 1)), 6),
    (dict(dom=(0.0, 1.0, 0.0, 1.0, 2.0, 1.0)), 7),
    (dict(n=(8, 8, 2)), 8),
    (dict(bc=(1, 1, 7)), 9),
    (dict(ldimf=8), 10),
    (dict(mdimf=8), 11),
    (dict(n=(9, 8, 8)), 13),
]


@pytest.mark.parametrize("change,code", BAD)
def test_ierror_matches_oracle(change, code):
    """Invalid arguments are rejected on the host with the oracle's IERROR."""
    n = change.get("n", (8, 8, 8))
    bc = change.get("bc", (1, 1, 1))
    dom = change.get("dom", (0.0, 1.0) * 3)
    ld = change.get("ldimf", n[0] + 1)
    md = change.get("mdimf", n[1] + 1)
    f = np.zeros((n[2] + 1) * max(md, 1) * max(ld, 1))
    _, ierr = C.hw3crt(f, n, bc, -1.0, dom=dom, ldimf=ld, mdimf=md)
    assert ierr == code
    if "ldimf" not in change and "mdimf" not in change:
        _, oerr, _ = O.solver_3d(np.zeros((n[2] + 1, n[1] + 1, n[0] + 1)), n, bc, -1.0, dom=dom)
        assert oerr == code


def test_unsupported_size_reported():
    """Odd cell counts are rejected (the GPU transforms take any even count;
    FISHPACK takes odd ones too)."""
    f = np.zeros(22 * 9 * 9)
    _, ierr = C.hw3crt(f, (21, 8, 8), (1, 1, 1), -1.0)
    assert ierr == 13
    assert b"must be even" in C.lib().chz_last_error()


def test_drivers_built_and_print_usage():
    for prog, usage in [("testgpu", "N xbc ybc zbc"), ("testcpu", "N xbc ybc zbc"),
                        ("testgpurectangle", "Nx Ny Nz xbc ybc zbc"), ("testcpurectangle", "Nx Ny Nz xbc ybc zbc")]:
        r = subprocess.run([os.path.join(BIN, prog)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 255 and usage in r.stdout, (prog, r.stdout)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="no GPU")
@pytest.mark.parametrize("n", [(6, 8, 10), (16, 12, 8)])
def test_hw3crt_vs_oracle_every_code(n):
    rng = np.random.default_rng(7)
    for bc in itertools.product(range(5), repeat=3):
        singular = all(b in (0, 3) for b in bc)
        for lam in ((-2.3, 0.0) if singular else (-2.3,)):
            f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
            bcs = faces(rng, n, bc)
            want, oerr, opert = O.solver_3d(f, n, bc, lam, bcs)
            got = f.copy()
            pert, ierr = C.hw3crt(got, n, bc, lam, bcs)
            assert ierr == oerr == 0, (bc, lam, ierr, oerr)
            scale = max(np.max(np.abs(want)), 1e-300)
            assert np.max(np.abs(got - want)) / scale < TOL, (bc, lam)
            assert abs(pert - opert) <= TOL * max(abs(opert), np.max(np.abs(f))), (bc, lam, pert, opert)
            if not singular:
                assert pert == 0.0


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="no GPU")
def test_hw3crt_leading_dimensions_and_lambda_warning():
    """F(LDIMF, MDIMF, N+1) and the face arrays with leading dimensions larger than
    the node counts: same solution as the packed call, padding untouched;
    ELMBDA > 0 returns IERROR 12 with the solution computed."""
    rng = np.random.default_rng(8)
    n, bc, lam = (8, 6, 10), (3, 2, 4), 0.7
    NX, NY, NZ = n
    ld, md = NX + 4, NY + 3
    f = rng.standard_normal((NZ + 1, NY + 1, NX + 1))
    bcs = faces(rng, n, bc)
    want, oerr, _ = O.solver_3d(f, n, bc, lam, bcs)
    assert oerr == 12
    packed = f.copy()
    _, ierr = C.hw3crt(packed, n, bc, lam, bcs)
    assert ierr == 12
    assert np.max(np.abs(packed - want)) / np.max(np.abs(want)) < TOL
    big = np.full((NZ + 1, md, ld), 7.25)
    big[:, :NY + 1, :NX + 1] = f
    # faces in FISHPACK's padded layouts: BDXS(MDIMF, N+1), BDYS(LDIMF, N+1), BDZS(LDIMF, M+1)
    shapes = [(NZ + 1, NY + 1, md), (NZ + 1, NX + 1, ld), (NY + 1, NX + 1, ld)]
    pb = []
    for i, b in enumerate(bcs):
        if b is None:
            pb.append(None)
            continue
        rows, cols, lead = shapes[i // 2]
        a = np.full((rows, lead), -3.5)
        a[:, :cols] = b.reshape(rows, cols)
        pb.append(a)
    _, ierr = C.hw3crt(big.reshape(-1), n, bc, lam, pb, ldimf=ld, mdimf=md)
    assert ierr == 12
    assert np.array_equal(big[:, :NY + 1, :NX + 1], packed)
    assert np.all(big[:, NY + 1:, :] == 7.25) and np.all(big[:, :, NX + 1:] == 7.25)


def manufactured(n, bc, lam=-1.0):
    """The drivers' problem: u = 1/3 sin(2 pi x) sin(2 pi y) sin(2 pi z) on [0,1]^3
    (the reference's ufunc.h case 1), f = laplace u + lam u, Dirichlet values
    and Neumann derivatives (along +axis) on the faces each code needs."""
    NX, NY, NZ = n
    x = np.arange(NX + 1) / NX
    y = np.arange(NY + 1) / NY
    z = np.arange(NZ + 1) / NZ
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    s = lambda t: np.sin(2 * np.pi * t)
    c = lambda t: np.cos(2 * np.pi * t)
    u = s(X) * s(Y) * s(Z) / 3
    f = -4 * np.pi ** 2 * s(X) * s(Y) * s(Z) + lam * u
    if bc[0] in (1, 2): f[:, :, 0] = u[:, :, 0]
    if bc[0] in (1, 4): f[:, :, NX] = u[:, :, NX]
    if bc[1] in (1, 2): f[:, 0, :] = u[:, 0, :]
    if bc[1] in (1, 4): f[:, NY, :] = u[:, NY, :]
    if bc[2] in (1, 2): f[0, :, :] = u[0, :, :]
    if bc[2] in (1, 4): f[NZ, :, :] = u[NZ, :, :]
    d = 2 * np.pi / 3
    dx = d * c(X) * s(Y) * s(Z)
    dy = d * s(X) * c(Y) * s(Z)
    dz = d * s(X) * s(Y) * c(Z)
    lo = lambda b: b in (3, 4)
    hi = lambda b: b in (2, 3)
    bcs = [dx[:, :, 0].copy() if lo(bc[0]) else None, dx[:, :, NX].copy() if hi(bc[0]) else None,
           dy[:, 0, :].copy() if lo(bc[1]) else None, dy[:, NY, :].copy() if hi(bc[1]) else None,
           dz[0, :, :].copy() if lo(bc[2]) else None, dz[NZ, :, :].copy() if hi(bc[2]) else None]
    return u, f, [None if b is None else b.reshape(-1) for b in bcs]


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="no GPU")
@pytest.mark.parametrize("prog,args", [("testgpu", [16, 1, 2, 0]), ("testcpu", [16, 3, 4, 1]),
                                       ("testgpurectangle", [16, 12, 8, 2, 0, 3]),
                                       ("testcpurectangle", [8, 16, 12, 4, 1, 2])])
def test_drivers_max_diff(prog, args):
    """The driver's printed max |u_h - u| equals the oracle's for the same discrete problem."""
    n = tuple(args[:3]) if "rectangle" in prog else (args[0],) * 3
    bc = tuple(args[-3:])
    r = subprocess.run([os.path.join(BIN, prog)] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert re.search(r"^\|+ %s: \t\d+\.\d{6}$" % prog, r.stdout, re.M), r.stdout
    if prog.startswith("testcpu"):
        assert "State code : 0" in r.stdout
    got = float(re.search(r"max diff : (\d+\.\d{14})", r.stdout).group(1))
    u, f, bcs = manufactured(n, bc)
    uh = O.helmholtz(f, n, bc, -1.0, bcs)
    want = float(np.max(np.abs(uh - u)))
    assert abs(got - want) <= 2e-14 + 1e-9 * want, (got, want)
