"""CPU oracle for the cuhelmholtz_amd HIP library.

TEST INFRASTRUCTURE ONLY: loaded by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` as the checker.  It restates the
reference's transform and solver in plain C (``chz_oracle.c``, every routine
citing the reference file:line it follows) and adds a CPU solver with the
FISHPACK ``hw3crt`` contract.  No product code imports it.

Pinning (tests/test_oracle.py): closed-form O(N^2) sums, scipy.fft golden
vectors, round-trip identities, a sparse direct solve of the 7-point system,
the residual invariant and O(h^2) manufactured solutions.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libchz_oracle.so")

DCT_1, DCT_2, DCT_3, DST_1, DST_2, DST_3, DFT_R2C, DFT_C2R = range(8)
KIND_NAMES = ["DCT_1", "DCT_2", "DCT_3", "DST_1", "DST_2", "DST_3", "DFT_R2C", "DFT_C2R"]


def build():
    """Compile the oracle C restatement with gcc (via oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_line.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        L.oracle_transform.argtypes = [dp] + [ctypes.c_int] * 6
        L.oracle_helmholtz.argtypes = ([ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp, dp] * 3
                                       + [ctypes.c_double, dp])
        L.oracle_helmholtz19.argtypes = L.oracle_helmholtz.argtypes
        L.oracle_bc_fold.argtypes = ([ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp, dp] * 3
                                     + [dp])
        L.solver_3d.argtypes = ([ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp, dp] * 3
                                + [ctypes.c_double, dp, dp])
        L.oracle_residual.argtypes = ([ctypes.c_int] * 6 + [ctypes.c_double] * 4 + [dp, dp, dp])
        L.oracle_residual.restype = ctypes.c_double
        L.hw3crt_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def line(kind, x):
    """One padded row (length N+2), returns a new array."""
    p = np.ascontiguousarray(x, dtype=np.float64).copy()
    N = p.size - 2
    if lib().oracle_line(kind, N, _p(p)) != 0:
        raise ValueError("bad kind or N")
    return p


def transform(vol, kx, ky, kz):
    """3D transform of a padded volume shaped (NZ+2, NY+2, NX+2) (x fastest)."""
    v = np.ascontiguousarray(vol, dtype=np.float64).copy()
    NZ, NY, NX = (s - 2 for s in v.shape)
    if lib().oracle_transform(_p(v), NX, NY, NZ, kx, ky, kz) != 0:
        raise ValueError("bad transform arguments")
    return v


def _bcargs(dom, n, bc, bcs):
    out = []
    for a in range(3):
        lo, hi = bcs[2 * a], bcs[2 * a + 1]
        out += [dom[2 * a], dom[2 * a + 1], n[a], bc[a],
                _p(None if lo is None else np.ascontiguousarray(lo, dtype=np.float64)),
                _p(None if hi is None else np.ascontiguousarray(hi, dtype=np.float64))]
    return out


def helmholtz(f, n, bc, lam, bcs=(None,) * 6, dom=(0.0, 1.0) * 3, stencil=7):
    """Restated cuda_helmholtz_solver on a node array f shaped (NZ+1, NY+1, NX+1).
    bcs = (bcl, bcr, bcb, bct, bcf, bce) face arrays or None. Returns u.
    stencil=19: the 19-point compact operator (oracle_helmholtz19)."""
    u = np.ascontiguousarray(f, dtype=np.float64).copy()
    keep = [None if b is None else np.ascontiguousarray(b, dtype=np.float64) for b in bcs]
    fn = lib().oracle_helmholtz if stencil == 7 else lib().oracle_helmholtz19
    if fn(*_bcargs(dom, n, bc, keep), float(lam), _p(u)) != 0:
        raise ValueError("bad solver arguments")
    return u


def bc_fold(f, n, bc, bcs=(None,) * 6, dom=(0.0, 1.0) * 3):
    u = np.ascontiguousarray(f, dtype=np.float64).copy()
    keep = [None if b is None else np.ascontiguousarray(b, dtype=np.float64) for b in bcs]
    if lib().oracle_bc_fold(*_bcargs(dom, n, bc, keep), _p(u)) != 0:
        raise ValueError("bad arguments")
    return u


def solver_3d(f, n, bc, lam, bcs=(None,) * 6, dom=(0.0, 1.0) * 3, threads=1):
    """CPU solver with the hw3crt contract. Returns (u, ierror, pertrb)."""
    u = np.ascontiguousarray(f, dtype=np.float64).copy()
    keep = [None if b is None else np.ascontiguousarray(b, dtype=np.float64) for b in bcs]
    pert = np.zeros(1)
    lib().hw3crt_set_threads(int(threads))
    ierr = lib().solver_3d(*_bcargs(dom, n, bc, keep), float(lam), _p(u), _p(pert))
    return u, ierr, float(pert[0])


def residual(u, ffold, n, bc, lam, dom=(0.0, 1.0) * 3):
    """max |ffold - (Delta_h u + lam u)| over the unknown box."""
    h = [(dom[2 * a + 1] - dom[2 * a]) / n[a] for a in range(3)]
    u = np.ascontiguousarray(u, dtype=np.float64)
    ff = np.ascontiguousarray(ffold, dtype=np.float64)
    r = np.zeros_like(u)
    m = lib().oracle_residual(n[0], bc[0], n[1], bc[1], n[2], bc[2], h[0], h[1], h[2], float(lam),
                              _p(u), _p(ff), _p(r))
    return m, r


# ---------------------------------------------------------------------------
# Windows and bc tables shared by the tests (reference conventions)
# ---------------------------------------------------------------------------
def in_window(kind, N):
    """Input slots a kind reads (cudasymmfft/*funcinplace.cu preOps)."""
    return {DST_1: (1, N - 1), DCT_1: (0, N), DCT_2: (0, N - 1), DST_2: (1, N),
            DCT_3: (0, N - 1), DST_3: (1, N), DFT_R2C: (0, N - 1), DFT_C2R: (0, N + 1)}[kind]


def out_window(kind, N):
    """Output slots a kind defines (cudasymmfft/*funcinplace.cu postOps)."""
    return {DST_1: (1, N - 1), DCT_1: (0, N), DCT_2: (0, N - 1), DST_2: (1, N),
            DCT_3: (0, N - 1), DST_3: (1, N), DFT_R2C: (0, N + 1), DFT_C2R: (0, N - 1)}[kind]


BC_INFO = {  # code -> (start offset, end offset from N, fwd kind, inv kind), cuda_helmholtz_solver.cu:78-172
    0: (0, -1, DFT_R2C, DFT_C2R),
    1: (1, -1, DST_1, DST_1),
    2: (1, 0, DST_3, DST_2),
    3: (0, 0, DCT_1, DCT_1),
    4: (0, -1, DCT_3, DCT_2),
}


def unknown_window(bc, N):
    s, e, _, _ = BC_INFO[bc]
    return s, N + e
