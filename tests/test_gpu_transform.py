"""GPU parity of do_transform (HIP, through the C ABI) against the oracle.

Bar: 1e-12 relative L-infinity (max |gpu - ref| / max |ref|) on the valid
output windows, FP64.  Inputs fill only the product of the per-axis input
windows (pads 0), as the reference's kinds require (e.g. DST_3 needs p[0] = 0,
cudasymmfft/dst3funcinplace.cu:49-67).  Full-size shapes (512^3,
512x384x256) are checked through size-independent properties: separable
(rank-1) inputs against outer products of 1-D oracle rows, and forward ->
inverse round trips.
"""
import itertools

import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not torch.cuda.is_available(), reason="no GPU")]

import cuhelmholtz_amd as C  # noqa: E402

TOL = 1e-12
FWD7 = [O.DFT_R2C, O.DCT_1, O.DCT_2, O.DCT_3, O.DST_1, O.DST_2, O.DST_3]


def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def rand_volume(rng, kinds, n):
    NX, NY, NZ = n
    v = np.zeros((NZ + 2, NY + 2, NX + 2))
    w = [O.in_window(k, N) for k, N in zip(kinds, n)]
    sl = (slice(w[2][0], w[2][1] + 1), slice(w[1][0], w[1][1] + 1), slice(w[0][0], w[0][1] + 1))
    v[sl] = rng.uniform(-1, 1, v[sl].shape)
    # C2R inputs: imaginary parts of DC / Nyquist are zero along that axis
    if kinds[0] == O.DFT_C2R:
        v[:, :, 1] = v[:, :, NX + 1] = 0
    if kinds[1] == O.DFT_C2R:
        v[:, 1, :] = v[:, NY + 1, :] = 0
    if kinds[2] == O.DFT_C2R:
        v[1, :, :] = v[NZ + 1, :, :] = 0
    return v


def out_slices(kinds, n):
    w = [O.out_window(k, N) for k, N in zip(kinds, n)]
    return (slice(w[2][0], w[2][1] + 1), slice(w[1][0], w[1][1] + 1), slice(w[0][0], w[0][1] + 1))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def gpu_transform(v, kinds, inplace=True, dtype=torch.float64):
    x = torch.from_numpy(v).to(dev(), dtype)
    out = None if inplace else torch.full_like(x, float("nan"))
    y = C.do_transform(x, out, kinds)
    torch.cuda.synchronize()
    return y.double().cpu().numpy()


@pytest.mark.parametrize("n", [(12, 8, 6), (8, 16, 4)])
def test_all_kind_triples_small(n):
    rng = np.random.default_rng(0)
    worst = 0.0
    for kinds in itertools.product(range(8), repeat=3):
        v = rand_volume(rng, kinds, n)
        ref = O.transform(v, *kinds)
        sl = out_slices(kinds, n)
        for inplace in (True, False):
            got = gpu_transform(v, kinds, inplace)
            e = rel(got[sl], ref[sl])
            worst = max(worst, e)
            assert e < TOL, (kinds, inplace, e)
    print("worst rel err", worst)


@pytest.mark.parametrize("N", [2, 4, 6, 8, 12, 16, 18, 24, 32, 36, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024,
                               14, 22, 26, 28, 34, 42, 98, 700, 442, 514, 1018])
def test_axis_lengths_all_kinds(N):
    """Every kind along every axis for 2,3,5-smooth lengths (all specialised
    radix plans) and for lengths whose N/2 has other prime factors (direct
    stages: 7, 11, 13, 17, 7^2, 2 5^2 7, 13 17, 257 and 509 -- the last two
    larger than a plan byte, run as one direct stage of the remaining factor)."""
    rng = np.random.default_rng(N)
    for axis in range(3):
        for kind in range(8):
            n = [6, 4, 8]
            n[axis] = N
            kinds = [O.DCT_2, O.DST_1, O.DCT_1]
            kinds[axis] = kind
            v = rand_volume(rng, kinds, n)
            ref = O.transform(v, *kinds)
            got = gpu_transform(v, kinds, inplace=(axis % 2 == 0))
            sl = out_slices(kinds, n)
            assert rel(got[sl], ref[sl]) < TOL, (N, axis, kind)


def test_cube_128_sampled():
    rng = np.random.default_rng(0)
    n = (128, 128, 128)
    triples = [tuple(rng.choice(FWD7, 3)) for _ in range(6)] + [(O.DST_3, O.DCT_3, O.DFT_R2C), (O.DCT_1, O.DST_1, O.DST_2)]
    for t, kinds in enumerate(triples):
        v = rand_volume(rng, kinds, n)
        ref = O.transform(v, *kinds)
        got = gpu_transform(v, kinds, inplace=(t % 2 == 0))
        sl = out_slices(kinds, n)
        assert rel(got[sl], ref[sl]) < TOL, kinds


def line_of(rng, kind, N):
    x = np.zeros(N + 2)
    lo, hi = O.in_window(kind, N)
    x[lo:hi + 1] = rng.uniform(-1, 1, hi - lo + 1)
    return x


@pytest.mark.parametrize("n,inplace", [((512, 512, 512), True), ((512, 384, 256), False)])
def test_full_size_separable(n, inplace):
    """Rank-1 input a_i b_j c_k: the transform is the outer product of the three
    1-D oracle rows, at the full benchmark shapes."""
    rng = np.random.default_rng(1)
    NX, NY, NZ = n
    for kinds in [(O.DCT_1, O.DST_3, O.DFT_R2C), (O.DST_1, O.DCT_2, O.DCT_3), (O.DFT_R2C, O.DST_2, O.DCT_1)]:
        a, b, c = (line_of(rng, k, N) for k, N in zip(kinds, n))
        x = torch.from_numpy(c).to(dev())[:, None, None] * torch.from_numpy(b).to(dev())[None, :, None] \
            * torch.from_numpy(a).to(dev())[None, None, :]
        x = x.contiguous()
        out = None if inplace else torch.empty_like(x)
        y = C.do_transform(x, out, kinds)
        ta, tb, tc = (torch.from_numpy(O.line(k, l)).to(dev()) for k, l in zip(kinds, (a, b, c)))
        sl = out_slices(kinds, n)
        ref = (tc[:, None, None] * tb[None, :, None] * ta[None, None, :])[sl]
        err = ((y[sl] - ref).abs().max() / ref.abs().max()).item()
        assert err < TOL, (kinds, err)
        del x, y, ref
        torch.cuda.empty_cache()


@pytest.mark.parametrize("n", [(512, 384, 256), (512, 512, 512)])
def test_full_size_roundtrip(n):
    """forward then inverse kinds give (prod N_a/2, DFT axes N_a) x input."""
    rng = np.random.default_rng(2)
    NX, NY, NZ = n
    fwd = (O.DCT_3, O.DST_3, O.DFT_R2C)
    inv = (O.DCT_2, O.DST_2, O.DFT_C2R)
    v = torch.zeros((NZ + 2, NY + 2, NX + 2), dtype=torch.float64, device=dev())
    g = torch.Generator(device=dev()).manual_seed(3)
    v[0:NZ, 1:NY + 1, 0:NX] = torch.rand((NZ, NY, NX), generator=g, device=dev(), dtype=torch.float64) * 2 - 1
    w = torch.empty_like(v)
    C.do_transform(v, w, fwd)
    C.do_transform(w, None, inv)
    scale = (NX / 2) * (NY / 2) * NZ
    err = ((w[0:NZ, 1:NY + 1, 0:NX] / scale - v[0:NZ, 1:NY + 1, 0:NX]).abs().max()).item()
    assert err < TOL, err


def test_f32_path_matches_oracle():
    rng = np.random.default_rng(4)
    n = (64, 48, 32)
    for kinds in [(O.DCT_1, O.DST_3, O.DFT_R2C), (O.DST_2, O.DCT_3, O.DST_1), (O.DFT_C2R, O.DCT_2, O.DST_3)]:
        v = rand_volume(rng, kinds, n)
        ref = O.transform(v, *kinds)
        got = gpu_transform(v, kinds, inplace=True, dtype=torch.float32)
        sl = out_slices(kinds, n)
        assert rel(got[sl], ref[sl]) < 2e-6, kinds


@pytest.mark.parametrize("n,kinds", [((512, 256, 16), (O.DCT_2, O.DST_1, O.DFT_R2C)),
                                     ((16, 384, 512), (O.DST_3, O.DCT_1, O.DCT_3)),
                                     ((256, 512, 8), (O.DFT_R2C, O.DST_2, O.DCT_2))])
def test_f32_path_large_lines(n, kinds):
    """FP32 transforms on the lean-I/O shapes of N = 256, 384, 512 (rows, columns, z columns)"""
    rng = np.random.default_rng(5)
    v = rand_volume(rng, kinds, n)
    ref = O.transform(v, *kinds)
    got = gpu_transform(v, kinds, inplace=True, dtype=torch.float32)
    sl = out_slices(kinds, n)
    assert rel(got[sl], ref[sl]) < 5e-6, kinds


def test_errors_are_reported():
    x = torch.zeros((8, 8, 9), dtype=torch.float64, device=dev())   # NX = 7 (odd)
    with pytest.raises(C.CHZError, match="even"):
        C.do_transform(x, None, (O.DCT_2, O.DCT_2, O.DCT_2))
    y = torch.zeros((8, 8, 8), dtype=torch.float64, device=dev())
    with pytest.raises(C.CHZError, match="kind"):
        C.do_transform(y, None, (9, O.DCT_2, O.DCT_2))
    # NX = 16400: more points per line than a workgroup holds (512 threads x 16)
    z = torch.zeros((4, 4, 16402), dtype=torch.float64, device=dev())
    with pytest.raises(C.CHZError, match="no launch configuration"):
        C.do_transform(z, None, (O.DCT_2, O.DCT_2, O.DCT_2))


def test_cpp_dropin_do_transform():
    """The C++ drop-in (reference signature, mangled name) computes the same result."""
    import ctypes
    L = C.lib()
    fn = getattr(L, "_Z12do_transformPdS_iii14transform_kindS0_S0_")
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 6
    rng = np.random.default_rng(5)
    n = (16, 12, 8)
    kinds = (O.DST_3, O.DCT_1, O.DFT_R2C)
    v = rand_volume(rng, kinds, n)
    ref = O.transform(v, *kinds)
    x = torch.from_numpy(v).to(dev())
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    fn(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), *n, *kinds)
    sl = out_slices(kinds, n)
    assert rel(y.cpu().numpy()[sl], ref[sl]) < TOL


def test_cpp_dropin_cubic_and_nocubic():
    """do_transform_cubic (in place) and do_transform_nocubic (result in d_out),
    reference cudasymmfft/funcinterface.h:14-29, against the oracle."""
    import ctypes
    L = C.lib()
    cubic = getattr(L, "_Z18do_transform_cubicPdiii14transform_kindS0_S0_")
    cubic.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6
    nocubic = getattr(L, "_Z20do_transform_nocubicPdS_iii14transform_kindS0_S0_")
    nocubic.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 6
    rng = np.random.default_rng(6)
    for n, kinds in [((16, 16, 16), (O.DCT_2, O.DST_1, O.DFT_R2C)), ((14, 14, 14), (O.DST_2, O.DCT_3, O.DCT_1))]:
        v = rand_volume(rng, kinds, n)
        ref = O.transform(v, *kinds)
        sl = out_slices(kinds, n)
        x = torch.from_numpy(v).to(dev())
        torch.cuda.synchronize()
        cubic(ctypes.c_void_p(x.data_ptr()), *n, *kinds)
        assert rel(x.cpu().numpy()[sl], ref[sl]) < TOL, (n, "cubic")
    n, kinds = (16, 12, 8), (O.DST_3, O.DCT_1, O.DFT_R2C)
    v = rand_volume(rng, kinds, n)
    ref = O.transform(v, *kinds)
    x = torch.from_numpy(v).to(dev())
    y = torch.full_like(x, float("nan"))
    torch.cuda.synchronize()
    nocubic(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), *n, *kinds)
    sl = out_slices(kinds, n)
    assert rel(y.cpu().numpy()[sl], ref[sl]) < TOL
