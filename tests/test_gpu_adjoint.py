"""GPU checks of the hand-written adjoint solve (chz_helmholtz_adjoint) and
the torch.autograd.Function over it.

- explicit FP64 adjoint: the transpose of the dense linear map (f, Neumann
  faces) -> u, built column by column from the oracle's restatement of
  cuda_helmholtz_solver at small N, applied to a random output gradient
  (1e-12 relative);
- d/dlambda: exact derivative from the oracle (du/dlambda = -A^-1 u on the
  unknown box, one more oracle solve), 1e-11 relative;
- adjoint identity <J x, y> = <x, J^T y> with the GPU forward and adjoint at
  128^3 (size-independent property, 1e-12);
- central finite differences (step 1e-6, FP64) on 32 random directions in
  (f, lambda) for L = 1/2 |u - u_target|^2 through autograd, on the
  differentiable workload (batch 16 of 128^3, codes (1,3,0), lambda = -2,
  f ~ N(0,1) seed 3, u_target from seed 4).
"""
import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not torch.cuda.is_available(), reason="no GPU")]

import cuhelmholtz_amd as C  # noqa: E402
from cuhelmholtz_amd.autograd import helmholtz  # noqa: E402


def dev():
    return torch.device("cuda:0")


def face_sizes(n):
    NX, NY, NZ = n
    return [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2


def used_faces(bc):
    return [bc[i // 2] in ((3, 4) if i % 2 == 0 else (2, 3)) for i in range(6)]


def dense_vjp(n, bc, lam, ubar, dom=(0.0, 1.0) * 3):
    """J^T ubar of the oracle solve, J built one unit input at a time."""
    NX, NY, NZ = n
    shape = (NZ + 1, NY + 1, NX + 1)
    sizes, used = face_sizes(n), used_faces(bc)
    zero_faces = [np.zeros(sizes[i]) if used[i] else None for i in range(6)]
    gf = np.zeros(shape)
    e = np.zeros(shape)
    for idx in range(e.size):
        e.flat[idx] = 1.0
        gf.flat[idx] = np.sum(ubar * O.helmholtz(e, n, bc, lam, zero_faces, dom))
        e.flat[idx] = 0.0
    gfaces = [None] * 6
    zf = np.zeros(shape)
    for i in range(6):
        if not used[i]:
            continue
        gfaces[i] = np.zeros(sizes[i])
        for t in range(sizes[i]):
            faces = [None if b is None else b.copy() for b in zero_faces]
            faces[i][t] = 1.0
            gfaces[i][t] = np.sum(ubar * O.helmholtz(zf, n, bc, lam, faces, dom))
    return gf, gfaces


CASES = [((8, 8, 8), (0, 3, 1), -1.5), ((12, 8, 10), (2, 4, 0), 0.3), ((8, 10, 8), (3, 3, 3), 0.0),
         ((10, 8, 12), (4, 2, 1), -7.0), ((8, 8, 10), (1, 1, 1), -0.5), ((8, 12, 8), (0, 0, 0), -2.0),
         ((10, 10, 8), (3, 0, 4), 0.0)]


@pytest.mark.parametrize("n,bc,lam", CASES)
def test_adjoint_matches_dense_transpose(n, bc, lam):
    rng = np.random.default_rng(hash((n, bc)) % 2 ** 31)
    NX, NY, NZ = n
    dom = (0.0, 1.0, -0.5, 1.5, 0.0, 2.0)
    ubar = rng.standard_normal((NZ + 1, NY + 1, NX + 1))
    want_f, want_faces = dense_vjp(n, bc, lam, ubar, dom)
    f = rng.standard_normal(ubar.shape)
    sizes, used = face_sizes(n), used_faces(bc)
    faces = [rng.standard_normal(sizes[i]) if used[i] else None for i in range(6)]
    u = O.helmholtz(f, n, bc, lam, faces, dom)
    gf, gfaces, glam = C.helmholtz_adjoint(torch.from_numpy(u).to(dev()), torch.from_numpy(ubar).to(dev()),
                                           n, bc, lam, dom, faces=True)
    gf = gf.cpu().numpy()
    scale = np.max(np.abs(want_f))
    assert np.max(np.abs(gf - want_f)) <= 1e-12 * scale
    for i in range(6):
        assert (gfaces[i] is None) == (want_faces[i] is None)
        if want_faces[i] is not None:
            assert np.max(np.abs(gfaces[i].cpu().numpy() - want_faces[i])) <= 1e-12 * scale
    # exact d/dlambda: du/dlambda = solve of -u on the unknown box, homogeneous boundary data
    if not (lam == 0.0 and all(b in (0, 3) for b in bc)):
        s = [(1 if b in (1, 2) else 0, N - 1 if b in (0, 1, 4) else N) for b, N in zip(bc, n)]
        rhs = np.zeros_like(u)
        box = (slice(s[2][0], s[2][1] + 1), slice(s[1][0], s[1][1] + 1), slice(s[0][0], s[0][1] + 1))
        rhs[box] = -u[box]
        zero_faces = [np.zeros(sizes[i]) if used[i] else None for i in range(6)]
        du = O.helmholtz(rhs, n, bc, lam, zero_faces, dom)
        want_lam = float(np.sum(ubar * du))
        assert abs(float(glam) - want_lam) <= 1e-11 * max(1.0, abs(want_lam)), (float(glam), want_lam)


@pytest.mark.parametrize("bc,lam", [((1, 3, 0), -2.0), ((2, 4, 3), 0.7), ((0, 0, 3), 0.0)])
def test_adjoint_identity_128(bc, lam):
    n = (128, 128, 128)
    g = torch.Generator(device=dev())
    g.manual_seed(5)
    shape = (129, 129, 129)
    sizes, used = face_sizes(n), used_faces(bc)
    x = torch.randn(shape, generator=g, dtype=torch.float64, device=dev())
    xf = [torch.randn(sizes[i], generator=g, dtype=torch.float64, device=dev()) if used[i] else None for i in range(6)]
    y = torch.randn(shape, generator=g, dtype=torch.float64, device=dev())
    Jx = x.clone()
    C.helmholtz_solve(Jx, n, bc, lam, xf)
    gf, gfaces, _ = C.helmholtz_adjoint(None, y, n, bc, lam, faces=True, want_lambda=False)
    lhs = float((Jx * y).sum())
    rhs = float((x * gf).sum()) + sum(float((a * b).sum()) for a, b in zip(xf, gfaces) if a is not None)
    assert abs(lhs - rhs) <= 1e-12 * (float(Jx.abs().sum() * y.abs().max()))


def config5(B=16, N=128):
    n = (N, N, N)
    g = torch.Generator(device=dev())
    g.manual_seed(3)
    f = torch.randn((B, N + 1, N + 1, N + 1), generator=g, dtype=torch.float64, device=dev())
    sizes = face_sizes(n)
    bcb = torch.zeros(sizes[2], dtype=torch.float64, device=dev())
    bct = torch.zeros(sizes[3], dtype=torch.float64, device=dev())
    g.manual_seed(4)
    ut = torch.randn((B, N + 1, N + 1, N + 1), generator=g, dtype=torch.float64, device=dev())
    return n, f, (None, None, bcb, bct, None, None), ut


@pytest.mark.parametrize("B,N", [(2, 16), (16, 128)])
def test_autograd_central_differences(B, N):
    n, f, bcs, ut = config5(B, N)
    bc = (1, 3, 0)
    lam = torch.tensor(-2.0, dtype=torch.float64, device=dev(), requires_grad=True)
    fr = f.clone().requires_grad_(True)
    u = helmholtz(fr, lam, bc, bcs)
    L = 0.5 * ((u - ut) ** 2).sum()
    L.backward()
    gf, gl = fr.grad, float(lam.grad)

    def loss(fv, lv):
        with torch.no_grad():
            return float(0.5 * ((helmholtz(fv, lv, bc, bcs) - ut) ** 2).sum())

    g = torch.Generator(device=dev())
    g.manual_seed(9)
    eps = 1e-6
    worst = 0.0
    for _ in range(32):
        df = torch.randn(f.shape, generator=g, dtype=torch.float64, device=dev())
        dl = float(torch.randn((), generator=g, dtype=torch.float64, device=dev()))
        nrm = float(torch.sqrt((df ** 2).sum() + dl ** 2))
        df /= nrm
        dl /= nrm
        fd = (loss(f + eps * df, -2.0 + eps * dl) - loss(f - eps * df, -2.0 - eps * dl)) / (2 * eps)
        an = float((gf * df).sum()) + gl * dl
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-3 * float(L.detach())))
    assert worst <= 1e-5, worst


def test_autograd_faces_and_batched_lambda():
    """face tensors and a per-problem lambda receive gradients consistent with the adjoint"""
    n = (12, 8, 10)
    bc = (3, 2, 0)
    B = 3
    rng = np.random.default_rng(8)
    f = torch.from_numpy(rng.standard_normal((B, 11, 9, 13))).to(dev()).requires_grad_(True)
    sizes = face_sizes(n)
    faces = [None] * 6
    faces[0] = torch.from_numpy(rng.standard_normal(sizes[0])).to(dev()).requires_grad_(True)   # shared
    faces[1] = torch.from_numpy(rng.standard_normal((B, sizes[1]))).to(dev()).requires_grad_(True)  # per problem
    faces[3] = torch.from_numpy(rng.standard_normal(sizes[3])).to(dev()).requires_grad_(True)
    lam = torch.tensor([-1.0, -3.0, 0.4], dtype=torch.float64, device=dev(), requires_grad=True)
    w = torch.from_numpy(rng.standard_normal((B, 11, 9, 13))).to(dev())
    u = helmholtz(f, lam, bc, faces)
    (u * w).sum().backward()
    for k in range(B):
        fk = [None if t is None else (t[k] if t.dim() == 2 else t).detach().cpu().numpy() for t in faces]
        uk = O.helmholtz(f[k].detach().cpu().numpy(), n, bc, float(lam[k]), fk)
        gfk, gfaces_k = dense_vjp(n, bc, float(lam[k]), w[k].cpu().numpy())
        assert np.max(np.abs(f.grad[k].cpu().numpy() - gfk)) <= 1e-12 * np.max(np.abs(gfk))
        assert np.max(np.abs(faces[1].grad[k].cpu().numpy() - gfaces_k[1])) <= 1e-12 * np.max(np.abs(gfk))
        assert np.allclose(u[k].detach().cpu().numpy(), uk, rtol=0, atol=1e-12 * np.max(np.abs(uk)))
    assert faces[0].grad.shape == faces[0].shape and faces[3].grad.shape == faces[3].shape
