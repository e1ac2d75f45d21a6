"""Generate the committed golden fixtures (run once, in a container with scipy).

The reference ships no golden vectors (its tests print an analytic error only,
cudahelmholtz/testgpu.cu:195-220), so the fixtures come from independent
libraries, never from the reference or the oracle:

* golden_1d.npz     -- every transform kind at N in {8, 12, 16, 24, 48} from
                       scipy.fft: the reference kinds are 0.5 x scipy's
                       norm="backward" DCT/DST types 1-3 on their windows
                       (cudasymmfft/dct3funcinplace.cu:82 "1/2 of fftw"),
                       DFT_R2C = numpy rfft interleaved, DFT_C2R = N * irfft.
* golden_3d.npz     -- separable 3D transforms of a 12x8x6 padded volume for
                       a set of kind triples, by applying the scipy 1-D forms
                       along x, then y, then z.
* golden_solver.npz -- solutions of the 7-point ghost-point Helmholtz system
                       (cuda_helmholtz_solver.cu semantics) from a sparse
                       direct solve (scipy.sparse.linalg.spsolve), for a seeded
                       sample of boundary-code triples at 4x6x8 cells.

Usage: python tests/golden/gen_golden.py
"""
import itertools
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from scipy import fft as sfft

HERE = os.path.dirname(os.path.abspath(__file__))
DCT_1, DCT_2, DCT_3, DST_1, DST_2, DST_3, DFT_R2C, DFT_C2R = range(8)

IN_WIN = {DST_1: (1, -1), DCT_1: (0, 0), DCT_2: (0, -1), DST_2: (1, 0), DCT_3: (0, -1), DST_3: (1, 0),
          DFT_R2C: (0, -1), DFT_C2R: (0, 1)}
OUT_WIN = {DST_1: (1, -1), DCT_1: (0, 0), DCT_2: (0, -1), DST_2: (1, 0), DCT_3: (0, -1), DST_3: (1, 0),
           DFT_R2C: (0, 1), DFT_C2R: (0, -1)}


def scipy_line(kind, x):
    """Reference kind of one padded row via scipy/numpy; zeros outside the output window."""
    N = x.size - 2
    out = np.zeros(N + 2)
    if kind == DST_1:
        out[1:N] = 0.5 * sfft.dst(x[1:N], 1)
    elif kind == DCT_1:
        out[0:N + 1] = 0.5 * sfft.dct(x[0:N + 1], 1)
    elif kind == DCT_2:
        out[0:N] = 0.5 * sfft.dct(x[0:N], 2)
    elif kind == DST_2:
        out[1:N + 1] = 0.5 * sfft.dst(x[1:N + 1], 2)
    elif kind == DCT_3:
        out[0:N] = 0.5 * sfft.dct(x[0:N], 3)
    elif kind == DST_3:
        out[1:N + 1] = 0.5 * sfft.dst(x[1:N + 1], 3)
    elif kind == DFT_R2C:
        X = np.fft.rfft(x[0:N])
        out[0::2] = X.real
        out[1::2] = X.imag
    elif kind == DFT_C2R:
        out[0:N] = np.fft.irfft(x[0::2] + 1j * x[1::2], N) * N
    return out


def random_line(rng, kind, N):
    x = np.zeros(N + 2)
    lo, hi = IN_WIN[kind]
    x[lo:N + hi + 1] = rng.uniform(-1, 1, N + hi + 1 - lo)
    if kind == DFT_C2R:
        x[1] = 0.0
        x[N + 1] = 0.0
    return x


def scipy_3d(vol, kinds):
    """vol shaped (NZ+2, NY+2, NX+2); x, then y, then z."""
    v = vol.copy()
    for axis, kind in enumerate(kinds):
        ax = 2 - axis
        v = np.apply_along_axis(lambda l: scipy_line(kind, l), ax, v)
    return v


def random_volume(rng, kinds, n):
    NX, NY, NZ = n
    v = np.zeros((NZ + 2, NY + 2, NX + 2))
    w = [(IN_WIN[k][0], N + IN_WIN[k][1]) for k, N in zip(kinds, n)]
    sl = (slice(w[2][0], w[2][1] + 1), slice(w[1][0], w[1][1] + 1), slice(w[0][0], w[0][1] + 1))
    v[sl] = rng.uniform(-1, 1, v[sl].shape)
    if kinds[0] == DFT_C2R:
        v[:, :, 1] = 0.0
        v[:, :, NX + 1] = 0.0
    return v


BC = {0: (0, -1), 1: (1, -1), 2: (1, 0), 3: (0, 0), 4: (0, -1)}


def fold(f, n, bc, bcs, h):
    """Boundary folding of cudahelmholtz/_cuda_helmholtz_.inl:70-227 (numpy)."""
    f = f.copy()
    for a in range(3):
        if bc[a] == 0:
            continue
        fa = np.moveaxis(f, 2 - a, 0)  # (N_a+1, ., .) view: remaining axes (c, b) with b<c
        b_ax, c_ax = [x for x in range(3) if x != a]
        s = [BC[bc[x]][0] for x in range(3)]
        e = [n[x] + BC[bc[x]][1] for x in range(3)]
        face_l = None if bcs[2 * a] is None else bcs[2 * a].reshape(n[c_ax] + 1, n[b_ax] + 1)
        face_r = None if bcs[2 * a + 1] is None else bcs[2 * a + 1].reshape(n[c_ax] + 1, n[b_ax] + 1)
        cs, bs = slice(s[c_ax], e[c_ax] + 1), slice(s[b_ax], e[b_ax] + 1)
        ih, ih2 = 1 / h[a], 1 / h[a] ** 2
        first, last = s[a], e[a]
        if bc[a] in (1, 2):
            fa[first, cs, bs] -= fa[0, cs, bs] * ih2
        else:
            fa[first, cs, bs] += 2 * face_l[cs, bs] * ih
        if bc[a] in (1, 4):
            fa[last, cs, bs] -= fa[n[a], cs, bs] * ih2
        else:
            fa[last, cs, bs] -= 2 * face_r[cs, bs] * ih
    return f


def sparse_solve(f, n, bc, lam, bcs, dom=(0.0, 1.0) * 3):
    h = [(dom[2 * a + 1] - dom[2 * a]) / n[a] for a in range(3)]
    ff = fold(f, n, bc, bcs, h)
    win = [(BC[bc[a]][0], n[a] + BC[bc[a]][1]) for a in range(3)]
    cnt = [w[1] - w[0] + 1 for w in win]

    def op(a):
        m = cnt[a]
        T = np.zeros((m, m))
        for i in range(m):
            T[i, i] = -2
            if i > 0:
                T[i, i - 1] = 1
            if i < m - 1:
                T[i, i + 1] = 1
        if bc[a] == 0:
            T[0, m - 1] += 1
            T[m - 1, 0] += 1
        if bc[a] in (3, 4):
            T[0, 1] = 2
        if bc[a] in (2, 3):
            T[m - 1, m - 2] = 2
        return sp.csr_matrix(T / h[a] ** 2)

    I = [sp.identity(c) for c in cnt]
    A = (sp.kron(I[2], sp.kron(I[1], op(0))) + sp.kron(I[2], sp.kron(op(1), I[0]))
         + sp.kron(op(2), sp.kron(I[1], I[0])) + lam * sp.identity(cnt[0] * cnt[1] * cnt[2]))
    sl = (slice(win[2][0], win[2][1] + 1), slice(win[1][0], win[1][1] + 1), slice(win[0][0], win[0][1] + 1))
    u = spla.spsolve(A.tocsc(), ff[sl].ravel()).reshape(ff[sl].shape)
    full = f.copy()
    full[sl] = u
    # periodic duplicate node N <- node 0 (x, then y, then z)
    if bc[0] == 0:
        full[:, :, n[0]] = full[:, :, 0]
    if bc[1] == 0:
        full[:, n[1], :] = full[:, 0, :]
    if bc[2] == 0:
        full[n[2], :, :] = full[0, :, :]
    return full


def face_arrays(rng, n, bc):
    NX, NY, NZ = n
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    out = []
    for a in range(3):
        need_l = bc[a] in (3, 4)
        need_r = bc[a] in (2, 3)
        out.append(rng.standard_normal(sizes[2 * a]) if need_l else None)
        out.append(rng.standard_normal(sizes[2 * a + 1]) if need_r else None)
    return out


def main():
    rng = np.random.default_rng(20240601)
    d1 = {}
    for N in (8, 12, 16, 24, 48):
        for kind in range(8):
            x = random_line(rng, kind, N)
            d1[f"x_{N}_{kind}"] = x
            d1[f"y_{N}_{kind}"] = scipy_line(kind, x)
    np.savez_compressed(os.path.join(HERE, "golden_1d.npz"), **d1)

    d3 = {}
    n = (12, 8, 6)
    triples = [(DCT_2, DST_1, DFT_R2C), (DST_3, DCT_1, DCT_3), (DFT_R2C, DST_2, DCT_2),
               (DCT_1, DFT_R2C, DST_3), (DST_1, DCT_3, DST_2), (DCT_3, DST_3, DCT_1),
               (DFT_C2R, DCT_2, DST_1), (DST_2, DFT_C2R, DFT_R2C), (DCT_2, DCT_2, DCT_2),
               (DFT_R2C, DFT_R2C, DFT_R2C)]
    for t, kinds in enumerate(triples):
        v = random_volume(rng, kinds, n)
        if kinds[1] == DFT_C2R:   # imaginary parts of DC / Nyquist along y are zero
            v[:, 1, :] = 0.0
            v[:, n[1] + 1, :] = 0.0
        d3[f"kinds_{t}"] = np.array(kinds)
        d3[f"x_{t}"] = v
        d3[f"y_{t}"] = scipy_3d(v, kinds)
    d3["n"] = np.array(n)
    d3["count"] = np.array(len(triples))
    np.savez_compressed(os.path.join(HERE, "golden_3d.npz"), **d3)

    ds = {}
    n = (4, 6, 8)
    combos = list(itertools.product(range(5), repeat=3))
    pick = [combos[i] for i in rng.choice(len(combos), 36, replace=False)]
    cases = [(bc, -1.7) for bc in pick] + [((0, 0, 0), 0.0), ((3, 3, 3), 0.0), ((1, 2, 3), 0.0), ((4, 0, 2), 0.5)]
    for t, (bc, lam) in enumerate(cases):
        f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
        bcs = face_arrays(rng, n, bc)
        singular = lam == 0.0 and all(b in (0, 3) for b in bc)
        ds[f"bc_{t}"] = np.array(bc)
        ds[f"lam_{t}"] = np.array(lam)
        ds[f"f_{t}"] = f
        for i, nm in enumerate(["bcl", "bcr", "bcb", "bct", "bcf", "bce"]):
            if bcs[i] is not None:
                ds[f"{nm}_{t}"] = bcs[i]
        if not singular:
            ds[f"u_{t}"] = sparse_solve(f, n, bc, lam, bcs)
    ds["n"] = np.array(n)
    ds["count"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "golden_solver.npz"), **ds)
    print("wrote golden_1d.npz, golden_3d.npz, golden_solver.npz")


if __name__ == "__main__":
    main()
