"""CPU tests that pin the oracle (oracle/chz_oracle.c) before it is trusted.

Each oracle function is checked against something other than its own
formula: O(N^2) closed forms of the transform kinds, scipy.fft golden vectors,
round-trip identities, a sparse direct solve of the 7-point system, the
residual invariant, and O(h^2) convergence on the manufactured solution of the
reference's own test driver shape (cudahelmholtz/testgpu.cu:57-122).
"""
import itertools
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KINDS = range(8)


def closed_form(kind, p):
    """Direct O(N^2) sums (definitions the reference's pre/post kernels implement)."""
    N = p.size - 2
    out = np.zeros(N + 2)
    j = np.arange(N + 1, dtype=float)
    if kind == O.DST_1:
        for k in range(1, N):
            out[k] = np.sum(p[1:N] * np.sin(np.pi * j[1:N] * k / N))
    elif kind == O.DCT_1:
        for k in range(N + 1):
            out[k] = 0.5 * p[0] + 0.5 * (-1) ** k * p[N] + np.sum(p[1:N] * np.cos(np.pi * j[1:N] * k / N))
    elif kind == O.DCT_2:
        for k in range(N):
            out[k] = np.sum(p[0:N] * np.cos(np.pi * (2 * j[0:N] + 1) * k / (2 * N)))
    elif kind == O.DST_2:
        for k in range(1, N + 1):
            out[k] = np.sum(p[1:N + 1] * np.sin(np.pi * (2 * j[1:N + 1] - 1) * k / (2 * N)))
    elif kind == O.DCT_3:
        for k in range(N):
            out[k] = 0.5 * p[0] + np.sum(p[1:N] * np.cos(np.pi * j[1:N] * (2 * k + 1) / (2 * N)))
    elif kind == O.DST_3:
        for k in range(1, N + 1):
            out[k] = 0.5 * (-1) ** (k - 1) * p[N] + np.sum(p[1:N] * np.sin(np.pi * j[1:N] * (2 * k - 1) / (2 * N)))
    elif kind == O.DFT_R2C:
        n = np.arange(N)
        for k in range(N // 2 + 1):
            out[2 * k] = np.sum(p[0:N] * np.cos(2 * np.pi * k * n / N))
            out[2 * k + 1] = -np.sum(p[0:N] * np.sin(2 * np.pi * k * n / N))
    elif kind == O.DFT_C2R:
        M = N // 2
        for t in range(N):
            s = p[0] + (-1) ** t * p[N]
            for k in range(1, M):
                a = 2 * np.pi * k * t / N
                s += 2 * (p[2 * k] * np.cos(a) - p[2 * k + 1] * np.sin(a))
            out[t] = s
    return out


def random_line(rng, kind, N):
    x = np.zeros(N + 2)
    lo, hi = O.in_window(kind, N)
    x[lo:hi + 1] = rng.uniform(-1, 1, hi - lo + 1)
    if kind == O.DFT_C2R:
        x[1] = x[N + 1] = 0.0
    return x


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("N", [4, 6, 8, 10, 12, 14, 16, 30, 48])
@pytest.mark.parametrize("kind", KINDS)
def test_line_closed_form(kind, N):
    rng = np.random.default_rng(N * 10 + kind)
    x = random_line(rng, kind, N)
    y = O.line(kind, x)
    lo, hi = O.out_window(kind, N)
    assert rel(y[lo:hi + 1], closed_form(kind, x)[lo:hi + 1]) < 1e-13


def test_line_golden():
    g = np.load(os.path.join(GOLD, "golden_1d.npz"))
    for N in (8, 12, 16, 24, 48):
        for kind in KINDS:
            x, want = g[f"x_{N}_{kind}"], g[f"y_{N}_{kind}"]
            lo, hi = O.out_window(kind, N)
            assert rel(O.line(kind, x)[lo:hi + 1], want[lo:hi + 1]) < 1e-13, (N, kind)


@pytest.mark.parametrize("N", [8, 12, 64, 96])
def test_line_roundtrips(N):
    """DST1^2 = DCT1^2 = DST2 o DST3 = DCT2 o DCT3 = (N/2) I, C2R o R2C = N I
    (the identities behind the solver's ratio, cuda_helmholtz_solver.cu:211-215)."""
    rng = np.random.default_rng(N)
    for fwd, inv, scale in [(O.DST_1, O.DST_1, N / 2), (O.DCT_1, O.DCT_1, N / 2), (O.DST_3, O.DST_2, N / 2),
                            (O.DCT_3, O.DCT_2, N / 2), (O.DST_2, O.DST_3, N / 2), (O.DCT_2, O.DCT_3, N / 2),
                            (O.DFT_R2C, O.DFT_C2R, N)]:
        x = random_line(rng, fwd, N)
        y = O.line(fwd, x)
        lo, hi = O.out_window(fwd, N)
        z = np.zeros(N + 2)
        z[lo:hi + 1] = y[lo:hi + 1]
        w = O.line(inv, z)
        lo, hi = O.in_window(fwd, N)
        assert rel(w[lo:hi + 1], scale * x[lo:hi + 1]) < 1e-13, (fwd, inv)


def test_3d_golden():
    g = np.load(os.path.join(GOLD, "golden_3d.npz"))
    NX, NY, NZ = (int(v) for v in g["n"])
    for t in range(int(g["count"])):
        kinds = [int(k) for k in g[f"kinds_{t}"]]
        y = O.transform(g[f"x_{t}"], *kinds)
        w = [O.out_window(k, N) for k, N in zip(kinds, (NX, NY, NZ))]
        sl = (slice(w[2][0], w[2][1] + 1), slice(w[1][0], w[1][1] + 1), slice(w[0][0], w[0][1] + 1))
        assert rel(y[sl], g[f"y_{t}"][sl]) < 1e-13, kinds


def golden_cases():
    g = np.load(os.path.join(GOLD, "golden_solver.npz"))
    n = tuple(int(v) for v in g["n"])
    for t in range(int(g["count"])):
        bcs = [g[f"{nm}_{t}"] if f"{nm}_{t}" in g else None for nm in ["bcl", "bcr", "bcb", "bct", "bcf", "bce"]]
        u = g[f"u_{t}"] if f"u_{t}" in g else None
        yield n, tuple(int(b) for b in g[f"bc_{t}"]), float(g[f"lam_{t}"]), g[f"f_{t}"], bcs, u


def test_solver_golden_sparse_direct():
    for n, bc, lam, f, bcs, u in golden_cases():
        if u is None:
            continue
        got = O.helmholtz(f, n, bc, lam, bcs)
        assert rel(got, u) < 1e-12, bc


def test_hw3crt_matches_golden():
    for n, bc, lam, f, bcs, u in golden_cases():
        got, ierr, _ = O.solver_3d(f, n, bc, lam, bcs)
        assert ierr in (0, 12)
        if u is not None:
            assert rel(got, u) < 1e-12, bc


def face_arrays(rng, n, bc):
    NX, NY, NZ = n
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    out = []
    for a in range(3):
        out.append(rng.standard_normal(sizes[2 * a]) if bc[a] in (3, 4) else None)
        out.append(rng.standard_normal(sizes[2 * a + 1]) if bc[a] in (2, 3) else None)
    return out


@pytest.mark.parametrize("n", [(6, 8, 10), (12, 4, 6)])
def test_residual_invariant_all_bc(n):
    """The restated solve satisfies the 7-point ghost-point system for all 125 bc triples."""
    rng = np.random.default_rng(7)
    for bc in itertools.product(range(5), repeat=3):
        lam = -2.3
        f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
        bcs = face_arrays(rng, n, bc)
        u = O.helmholtz(f, n, bc, lam, bcs)
        ff = O.bc_fold(f, n, bc, bcs)
        r, _ = O.residual(u, ff, n, bc, lam)
        anorm = 4 * sum(N * N for N in n) + abs(lam)
        assert r / (anorm * np.max(np.abs(u)) + np.max(np.abs(ff))) < 1e-14, bc


def test_hw3crt_matches_oracle_rect():
    rng = np.random.default_rng(11)
    n = (16, 12, 10)
    for bc in [(1, 2, 0), (3, 4, 0), (0, 0, 0), (2, 3, 1), (4, 1, 3), (0, 4, 2)]:
        f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
        bcs = face_arrays(rng, n, bc)
        dom = (0.0, 2.0, -1.0, 0.5, 0.25, 1.0)
        u = O.helmholtz(f, n, bc, -4.0, bcs, dom)
        v, ierr, _ = O.solver_3d(f, n, bc, -4.0, bcs, dom)
        assert ierr == 0
        assert rel(v, u) < 1e-12, bc


@pytest.mark.parametrize("bc", [(0, 0, 0), (3, 3, 3), (0, 3, 0), (3, 0, 3)])
def test_singular_case(bc):
    """lambda = 0 with only periodic/Neumann axes: the (0,0,0) mode is zeroed
    (cuda_helmholtz_solver.cu:328-330); hw3crt returns the same zero-mean
    solution and PERTRB = the weighted mean of the folded rhs."""
    rng = np.random.default_rng(3)
    n = (8, 6, 10)
    f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
    bcs = face_arrays(rng, n, bc)
    u = O.helmholtz(f, n, bc, 0.0, bcs)
    v, ierr, pert = O.solver_3d(f, n, bc, 0.0, bcs)
    assert ierr == 0
    assert rel(v, u) < 1e-12
    ff = O.bc_fold(f, n, bc, bcs)
    r, _ = O.residual(u, ff - pert, n, bc, 0.0)
    assert r < 1e-9 * np.max(np.abs(ff))


def test_ierror_codes():
    n = (8, 8, 8)
    f = np.zeros((9, 9, 9))
    assert O.solver_3d(f, n, (5, 1, 1), -1.0)[1] == 3
    assert O.solver_3d(f, n, (1, 7, 1), -1.0)[1] == 6
    assert O.solver_3d(f, n, (1, 1, -1), -1.0)[1] == 9
    assert O.solver_3d(f, n, (1, 1, 1), -1.0, dom=(1.0, 0.0, 0.0, 1.0, 0.0, 1.0))[1] == 1
    assert O.solver_3d(f, n, (1, 1, 1), 2.0)[1] == 12
    assert O.solver_3d(np.zeros((9, 9, 3)), (2, 8, 8), (1, 1, 1), -1.0)[1] == 2


def manufactured(n):
    """The manufactured-solution workload: u = sin(pi x) sin(pi y/2) cos(2 pi z), codes
    (1,2,0), lambda = -1 on [0,1]^3; f = Delta u + lambda u at the nodes with
    Dirichlet values written into the boundary nodes (testgpu.cu:57-122)."""
    NX, NY, NZ = n
    z, y, x = np.meshgrid(np.arange(NZ + 1) / NZ, np.arange(NY + 1) / NY, np.arange(NX + 1) / NX, indexing="ij")
    u = np.sin(np.pi * x) * np.sin(np.pi * y / 2) * np.cos(2 * np.pi * z)
    lam = -1.0
    f = (-np.pi ** 2 - np.pi ** 2 / 4 - 4 * np.pi ** 2 + lam) * u
    f[:, :, 0] = u[:, :, 0]
    f[:, :, NX] = u[:, :, NX]
    f[:, 0, :] = u[:, 0, :]
    bct = ((np.pi / 2) * np.sin(np.pi * x[:, NY, :]) * np.cos(np.pi / 2) * np.cos(2 * np.pi * z[:, NY, :])).ravel()
    return f, u, lam, [None, None, None, bct, None, None]


def test_manufactured_second_order():
    errs = []
    for N in (16, 32, 64):
        f, u, lam, bcs = manufactured((N, N, N))
        got = O.helmholtz(f, (N, N, N), (1, 2, 0), lam, bcs)
        errs.append(np.max(np.abs(got - u)))
        v, ierr, _ = O.solver_3d(f, (N, N, N), (1, 2, 0), lam, bcs)
        assert rel(v, got) < 1e-12
    assert 3.6 < errs[0] / errs[1] < 4.4 and 3.6 < errs[1] / errs[2] < 4.4, errs


def test_checks_catch_planted_mistakes():
    """The closed-form check must fail for plausible oracle mistakes: a
    swapped kind (DST-II for DST-III), a dropped endpoint term, a sign flip."""
    rng = np.random.default_rng(5)
    N = 12
    x = random_line(rng, O.DST_3, N)
    lo, hi = O.out_window(O.DST_3, N)
    good = closed_form(O.DST_3, x)[lo:hi + 1]
    wrong_kind = O.line(O.DST_2, x)[lo:hi + 1]
    assert rel(wrong_kind, good) > 1e-3
    x2 = x.copy()
    x2[N] = 0.0     # dropped 1/2 (-1)^(k-1) x_N term
    assert rel(O.line(O.DST_3, x2)[lo:hi + 1], good) > 1e-3
    assert rel(-O.line(O.DST_3, x)[lo:hi + 1], good) > 1e-3
    # a transposed operand in 3D: swapping x and y kinds changes the result
    g = np.load(os.path.join(GOLD, "golden_3d.npz"))
    k = [int(v) for v in g["kinds_1"]]
    y_sw = O.transform(g["x_1"], k[1], k[0], k[2])
    assert rel(y_sw[1:6, 1:8, 1:12], g["y_1"][1:6, 1:8, 1:12]) > 1e-3


# ---------------------------------------------------------------------------
# 19-point compact operator (the reference's unlaunched
# _nonpoisson_devide_19points_, cudahelmholtz/_cuda_helmholtz_.inl:255-274)
# ---------------------------------------------------------------------------
def _closure_d2(v, axis, code, h):
    """Second difference along `axis` of a grid function on the unknown box, with
    the code's homogeneous closure: periodic wrap (0), zero beyond Dirichlet ends
    (1, 2 low, 4 high), mirror ghost u_-1 = u_1 at Neumann ends (3, 2 high, 4 low)."""
    v = np.moveaxis(v, axis, 0)
    n = v.shape[0]
    if code == 0:
        lo, hi = v[-1:], v[:1]
    else:
        neu_lo, neu_hi = code in (3, 4), code in (2, 3)
        lo = v[1:2] if neu_lo else np.zeros_like(v[:1])
        hi = v[n - 2:n - 1] if neu_hi else np.zeros_like(v[:1])
    ext = np.concatenate([lo, v, hi], axis=0)
    d = (ext[2:] - 2 * ext[1:-1] + ext[:-2]) / h ** 2
    return np.moveaxis(d, 0, axis)


@pytest.mark.parametrize("bc", [(0, 0, 0), (1, 1, 1), (3, 3, 3), (0, 1, 3), (2, 4, 0), (4, 3, 2)])
def test_19point_matches_explicit_stencil(bc):
    """(L19 + lambda) u = B f on the unknown box, with L19 = sum D_a + sum_{a<b}
    (h_a^2 + h_b^2)/12 D_a D_b and B = I + sum h_a^2/12 D_a applied in physical
    space (homogeneous boundary data, unequal spacings): pins the spectral
    construction of oracle_helmholtz19 independently of its transforms."""
    rng = np.random.default_rng(19)
    n, dom, lam = (8, 6, 10), (0.0, 1.0, 0.0, 0.8, -0.5, 0.75), -1.7
    h = [(dom[2 * a + 1] - dom[2 * a]) / n[a] for a in range(3)]
    f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
    # homogeneous Dirichlet values, zero Neumann data
    if bc[0] in (1, 2): f[:, :, 0] = 0
    if bc[0] in (1, 4): f[:, :, -1] = 0
    if bc[1] in (1, 2): f[:, 0, :] = 0
    if bc[1] in (1, 4): f[:, -1, :] = 0
    if bc[2] in (1, 2): f[0] = 0
    if bc[2] in (1, 4): f[-1] = 0
    sizes = [(n[2] + 1) * (n[1] + 1)] * 2 + [(n[2] + 1) * (n[0] + 1)] * 2 + [(n[1] + 1) * (n[0] + 1)] * 2
    bcs = [np.zeros(sizes[i]) if bc[i // 2] in ((3, 4) if i % 2 == 0 else (2, 3)) else None for i in range(6)]
    u = O.helmholtz(f, n, bc, lam, bcs, dom, stencil=19)
    win = [O.unknown_window(bc[a], n[a]) for a in range(3)]
    box = (slice(win[2][0], win[2][1] + 1), slice(win[1][0], win[1][1] + 1), slice(win[0][0], win[0][1] + 1))
    ub, fb = u[box], f[box]
    ax = {0: 2, 1: 1, 2: 0}   # grid axis a -> numpy axis
    D = lambda v, a: _closure_d2(v, ax[a], bc[a], h[a])
    L = sum(D(ub, a) for a in range(3))
    for a, b in ((0, 1), (0, 2), (1, 2)):
        L = L + (h[a] ** 2 + h[b] ** 2) / 12 * D(D(ub, b), a)
    Bf = fb + sum(h[a] ** 2 / 12 * D(fb, a) for a in range(3))
    err = np.max(np.abs(L + lam * ub - Bf)) / np.max(np.abs(Bf))
    assert err < 1e-12, (bc, err)
    # and it is not the 7-point solution
    assert np.max(np.abs(u - O.helmholtz(f, n, bc, lam, bcs, dom))) > 1e-3 * np.max(np.abs(u))


def test_19point_fourth_order_poisson():
    """lambda = 0, periodic, smooth zero-mean u: the error falls as h^4."""
    errs = []
    for N in (16, 32, 64):
        x = np.arange(N + 1) / N
        Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
        u = np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) * np.sin(4 * np.pi * Z)
        f = -(4 + 4 + 16) * np.pi ** 2 * u
        uh = O.helmholtz(f, (N, N, N), (0, 0, 0), 0.0, stencil=19)
        errs.append(np.max(np.abs(uh - u)))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 13.5 < r1 < 16.5 and 15.2 < r2 < 16.5, errs   # 14.2, 15.5: h^4 (h^2 would give 4)
