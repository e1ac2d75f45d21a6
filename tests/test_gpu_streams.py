"""GPU checks of the library's state across calls: calls on different streams
that share the per-device workspaces, batched plans across chz_release, the
launch configurations across chz_release and a tuning change.

Every call is checked against the oracle's restatement of
cuda_helmholtz_solver (cudahelmholtz/cuda_helmholtz_solver.cu:33-380) or the
dense transpose it implies, at 1e-12 relative L-infinity.
"""
import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not torch.cuda.is_available(), reason="no GPU")]

import cuhelmholtz_amd as C  # noqa: E402

TOL = 1e-12


def dev():
    return torch.device("cuda:0")


def relnp(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def faces(rng, n, bc):
    NX, NY, NZ = n
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    out = []
    for i in range(6):
        need = bc[i // 2] in ((3, 4) if i % 2 == 0 else (2, 3))
        out.append(rng.standard_normal(sizes[i]) if need else None)
    return out


def problem(seed, n, bc, lam):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
    bcs = faces(rng, n, bc)
    return {"n": n, "bc": bc, "lam": lam, "f": f, "bcs": bcs,
            "want": O.helmholtz(f, n, bc, lam, bcs)}


def to_dev(p):
    return (torch.from_numpy(p["f"]).to(dev()),
            [None if b is None else torch.from_numpy(b).to(dev()) for b in p["bcs"]])


def test_two_streams_share_the_workspace():
    """Two solves of different geometries enqueued on two streams, no host
    sync in between: both use the FP64 workspace slot, and each must see only
    its own data.  The first is the larger one so the second cannot finish
    first by accident of size."""
    pa = problem(1, (96, 80, 64), (3, 4, 0), -4.0)
    pb = problem(2, (64, 48, 40), (1, 2, 3), 0.5)
    fa, ba = to_dev(pa)
    fb, bb = to_dev(pb)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for rep in range(3):
        ua, ub = fa.clone(), fb.clone()
        torch.cuda.synchronize()
        with torch.cuda.stream(s1):
            C.helmholtz_solve(ua, pa["n"], pa["bc"], pa["lam"], ba)
        with torch.cuda.stream(s2):
            C.helmholtz_solve(ub, pb["n"], pb["bc"], pb["lam"], bb)
        with torch.cuda.stream(s1):
            C.helmholtz_solve_mixed(ua.clone(), pa["n"], pa["bc"], pa["lam"], ba, refine=1)
        torch.cuda.synchronize()
        assert relnp(ua.cpu().numpy(), pa["want"]) <= TOL, rep
        assert relnp(ub.cpu().numpy(), pb["want"]) <= TOL, rep


def test_solve_and_adjoint_on_two_streams():
    """A solve on one stream and an adjoint (which runs a solve on the same
    workspace slot) on another, no host sync: the solve against the oracle,
    the adjoint against <J x, y> = <x, J^T y> with J the oracle's solve."""
    n, bc, lam = (64, 48, 32), (1, 3, 0), -2.0
    ps = problem(3, (80, 64, 48), (2, 1, 3), -1.0)
    fs, bs = to_dev(ps)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
    y = rng.standard_normal(x.shape)
    zero = [None, None, np.zeros((n[2] + 1) * (n[0] + 1)), np.zeros((n[2] + 1) * (n[0] + 1)), None, None]
    jx = O.helmholtz(x, n, bc, lam, zero)
    gy = torch.from_numpy(y).to(dev())
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    us = fs.clone()
    with torch.cuda.stream(s1):
        C.helmholtz_solve(us, ps["n"], ps["bc"], ps["lam"], bs)
    with torch.cuda.stream(s2):
        gf, _, _ = C.helmholtz_adjoint(None, gy, n, bc, lam, want_lambda=False)
    torch.cuda.synchronize()
    assert relnp(us.cpu().numpy(), ps["want"]) <= TOL
    lhs = float(np.sum(jx * y))
    rhs = float(np.sum(x * gf.cpu().numpy()))
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), 1.0), (lhs, rhs)


def test_batch_plan_survives_release():
    """A plan made by chz_batch_create owns its twiddle and eigenvalue tables:
    chz_release (which frees the library's caches) leaves it usable."""
    probs = [problem(10 + k, n, bc, lam) for k, (n, bc, lam) in enumerate(
        [((16, 12, 8), (1, 3, 0), -1.5), ((24, 16, 20), (0, 2, 4), -3.0), ((32, 32, 16), (3, 3, 1), 0.25)])]
    arrs = [to_dev(p) for p in probs]
    bs = C.BatchSolver([(f, p["n"], p["bc"], p["lam"], b) for p, (f, b) in zip(probs, arrs)])
    f0 = [f.clone() for f, _ in arrs]
    bs.solve()
    torch.cuda.synchronize()
    C.release()
    # other sizes after the release re-create the caches at new addresses
    other = problem(20, (40, 36, 30), (4, 0, 2), -2.0)
    fo, bo = to_dev(other)
    C.helmholtz_solve(fo, other["n"], other["bc"], other["lam"], bo)
    for (f, _), g in zip(arrs, f0):
        f.copy_(g)
    bs.solve()
    torch.cuda.synchronize()
    for p, (f, _) in zip(probs, arrs):
        assert relnp(f.cpu().numpy(), p["want"]) <= TOL, p["n"]
    assert relnp(fo.cpu().numpy(), other["want"]) <= TOL
    bs.close()


def test_rebind_on_another_stream_is_ordered():
    """Solve, rebind to new arrays and solve again on a second stream without
    a host sync: the second solve sees the new descriptors, the first result
    is untouched."""
    n, bc = (32, 24, 16), (1, 3, 0)
    pa = [problem(30 + k, n, bc, lam) for k, lam in enumerate((-1.0, -2.5))]
    pb = [problem(40 + k, n, bc, lam) for k, lam in enumerate((-0.5, -4.0))]
    da = [to_dev(p) for p in pa]
    db = [to_dev(p) for p in pb]
    bs = C.BatchSolver([(f, p["n"], p["bc"], p["lam"], b) for p, (f, b) in zip(pa, da)])
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    with torch.cuda.stream(s1):
        bs.solve()
    with torch.cuda.stream(s2):
        bs.rebind([(f, p["n"], p["bc"], p["lam"], b) for p, (f, b) in zip(pb, db)])
        bs.solve()
    torch.cuda.synchronize()
    for p, (f, _) in zip(pa + pb, da + db):
        assert relnp(f.cpu().numpy(), p["want"]) <= TOL, p["lam"]
    bs.close()


def test_configs_kept_across_release_and_cached_plan_rebuilt():
    """chz_release keeps the tuned launch configurations; a cached batched
    plan whose launch configuration was changed by a tuning call is rebuilt
    instead of failing."""
    before = [C.axis_config(512, 64, role=r) for r in range(5)]
    C.release()
    after = [C.axis_config(512, 64, role=r) for r in range(5)]
    assert before == after
    n, bc = (24, 20, 16), (3, 1, 0)
    p = problem(50, n, bc, -1.0)
    f, b = to_dev(p)
    u = f.clone().unsqueeze(0).contiguous()
    C.helmholtz_solve_batch(u, n, bc, -1.0, b)
    torch.cuda.synchronize()
    assert relnp(u[0].cpu().numpy(), p["want"]) <= TOL
    old = C.axis_config(24, 64, role=C.ROLE_XFWD)
    try:
        C.set_axis_config(24, 64, C.ROLE_XFWD, 4, 4, 8)
        u = f.clone().unsqueeze(0).contiguous()
        C.helmholtz_solve_batch(u, n, bc, -1.0, b)
        torch.cuda.synchronize()
        assert relnp(u[0].cpu().numpy(), p["want"]) <= TOL
    finally:
        C.set_axis_config(24, 64, C.ROLE_XFWD, old["ept"], old["tpl"], old["lines"])
