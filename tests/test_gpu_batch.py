"""GPU parity of the batched solve (one hipGraph, one persistent launch per
pass and size class, descriptor tables), through the C ABI.

Bar: every problem of a batch equals the single-problem solve of the same
problem (same kernels bodies and launch configurations) to 1e-13 relative
L-infinity, FP64 and FP32+refine; FP64 problems match the oracle's
restatement of cuda_helmholtz_solver on a sample to 1e-12; graph replays,
direct launches and repeated replays agree.
"""
import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not torch.cuda.is_available(), reason="no GPU")]

import cuhelmholtz_amd as C  # noqa: E402
from cuhelmholtz_amd import workloads as W  # noqa: E402

SAME = 1e-13


def dev():
    return torch.device("cuda:0")


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


def make(descs):
    """device rhs tensors, face tensors and host copies of a problem list"""
    out = []
    for d in descs:
        f, bcs = W.problem_arrays(d)
        ft = torch.from_numpy(f).to(dev())
        bt = [None if b is None else torch.from_numpy(b).to(dev()) for b in bcs]
        out.append({"d": d, "f0": ft.clone(), "f": ft, "bcs": bt, "f_np": f, "bcs_np": bcs})
    return out


def small_descs(count, seed, sizes=(8, 12, 16, 20, 24)):
    return W.batched_problems(count=count, seed=seed, sizes=sizes)


def single(p, prec=64, refine=0):
    u = p["f0"].clone()
    d = p["d"]
    if prec == 64:
        C.helmholtz_solve(u, d["n"], d["bc"], d["lam"], p["bcs"])
    else:
        C.helmholtz_solve_mixed(u, d["n"], d["bc"], d["lam"], p["bcs"], refine=refine)
    return u


def run_batch(probs, prec=64, refine=0, graph=True):
    bs = C.BatchSolver([(p["f"], p["d"]["n"], p["d"]["bc"], p["d"]["lam"], p["bcs"]) for p in probs], prec, refine)
    bs.solve(graph=graph)
    torch.cuda.synchronize()
    return bs


def test_batch_fp64_matches_single_small():
    probs = make(small_descs(40, 11))
    bs = run_batch(probs)
    info = bs.info()
    assert info["launches"] >= 5
    for k, p in enumerate(probs):
        e = rel(p["f"], single(p))
        assert e <= SAME, (k, p["d"], e)
    bs.close()


@pytest.mark.parametrize("sizes", [(8, 12, 16, 20, 24), (14, 22, 26, 34, 98)])
def test_batch_fp64_vs_oracle_sample(sizes):
    """A sample against the oracle; the second size set has N/2 with prime
    factors 7, 11, 13, 17 (the batched passes on direct FFT stages)."""
    probs = make(small_descs(20, 12, sizes))
    bs = run_batch(probs)
    for p in probs[::3]:
        d = p["d"]
        want = O.helmholtz(p["f_np"], d["n"], d["bc"], d["lam"], p["bcs_np"])
        got = p["f"].cpu().numpy()
        e = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
        assert e <= 1e-12, (d, e)
    bs.close()


def test_batch_graph_replays_and_eager_agree():
    probs = make(small_descs(24, 13))
    bs = C.BatchSolver([(p["f"], p["d"]["n"], p["d"]["bc"], p["d"]["lam"], p["bcs"]) for p in probs])
    want = [single(p) for p in probs]
    face0 = [[None if b is None else b.clone() for b in p["bcs"]] for p in probs]
    for rep in range(4):
        # new data in the bound arrays (rhs and Neumann faces scaled together:
        # the solve is linear in both), then solve: graph on even reps, direct
        # launches on odd ones
        for k, p in enumerate(probs):
            p["f"].copy_(p["f0"] * (1.0 + rep))
            for b, b0 in zip(p["bcs"], face0[k]):
                if b is not None:
                    b.copy_(b0 * (1.0 + rep))
        torch.cuda.synchronize()
        bs.solve(graph=rep % 2 == 0)
        torch.cuda.synchronize()
        for k, p in enumerate(probs):
            e = rel(p["f"], want[k] * (1.0 + rep))
            assert e <= SAME, (rep, k, e)
    bs.close()


@pytest.mark.parametrize("prec,refine", [(64, 0), (32, -3)])
def test_batch_graph_branches_replay_bitwise(prec, refine):
    """The graph runs the size classes of a step as parallel branches: on the
    benchmark's sizes (several classes per step; for FP32, per-problem
    refinement counts, so the steps hold uneven problem sets) every one of
    three replays, each after eager work and a synchronisation, equals the
    same launches issued directly on one stream, bit for bit."""
    probs = make(W.batched_problems(count=48, seed=5))
    bs = C.BatchSolver([(p["f"], p["d"]["n"], p["d"]["bc"], p["d"]["lam"], p["bcs"]) for p in probs], prec, refine)
    bs.solve(graph=False)
    torch.cuda.synchronize()
    want = [p["f"].clone() for p in probs]
    for rep in range(3):
        for p in probs:
            p["f"].copy_(p["f0"])
        torch.cuda.synchronize()
        bs.solve(graph=True)
        torch.cuda.synchronize()
        for k, p in enumerate(probs):
            assert torch.equal(p["f"], want[k]), (rep, k, p["d"])
    for k, p in enumerate(probs):
        assert rel(want[k], single(p)) <= (SAME if prec == 64 else 1e-10), (k, p["d"])
    bs.close()


def test_batch_benchmark_sizes_sample():
    """64 problems drawn like the 2048-problem workload (sizes 32..128)."""
    probs = make(W.batched_problems(count=64, seed=2))
    bs = run_batch(probs)
    for k, p in enumerate(probs):
        e = rel(p["f"], single(p))
        assert e <= SAME, (k, p["d"], e)
    bs.close()


@pytest.mark.parametrize("refine", [0, 2])
def test_batch_mixed_matches_single(refine):
    probs = make(small_descs(24, 14, sizes=(16, 24, 32, 48)))
    bs = run_batch(probs, prec=32, refine=refine)
    for k, p in enumerate(probs):
        e = rel(p["f"], single(p, 32, refine))
        assert e <= SAME, (k, p["d"], e)
        if refine == 2:
            assert rel(p["f"], single(p)) <= 1e-9, (k, p["d"])
    bs.close()


def test_batch_singular_fp64_and_periodic():
    descs = [{"n": (8, 12, 16), "bc": (0, 3, 0), "lam": 0.0, "seed": 5},
             {"n": (16, 8, 8), "bc": (0, 0, 0), "lam": -2.0, "seed": 6},
             {"n": (12, 12, 12), "bc": (3, 3, 3), "lam": 0.0, "seed": 7},
             {"n": (8, 16, 12), "bc": (1, 0, 2), "lam": 0.25, "seed": 8}]
    probs = make(descs)
    bs = run_batch(probs)
    for p in probs:
        d = p["d"]
        want = O.helmholtz(p["f_np"], d["n"], d["bc"], d["lam"], p["bcs_np"])
        got = p["f"].cpu().numpy()
        assert float(np.max(np.abs(got - want)) / np.max(np.abs(want))) <= 1e-12, d
        assert rel(p["f"], single(p)) <= SAME
    bs.close()


def test_batch_errors():
    probs = make([{"n": (8, 8, 8), "bc": (3, 3, 3), "lam": 0.0, "seed": 1}])
    p = probs[0]
    with pytest.raises(C.CHZError, match="singular"):
        C.BatchSolver([(p["f"], p["d"]["n"], p["d"]["bc"], 0.0, p["bcs"])], prec=32, refine=1)
    with pytest.raises(C.CHZError, match="problem 1"):
        C.BatchSolver([(p["f"], (8, 8, 8), (3, 3, 3), -1.0, p["bcs"]),
                       (p["f"], (8, 8, 8), (7, 3, 3), -1.0, p["bcs"])])
    with pytest.raises(C.CHZError):
        C.BatchSolver([(p["f"], (8, 8, 8), (3, 3, 3), -1.0, p["bcs"])], prec=64, refine=1)


def test_batch_rebind_new_arrays_and_lambdas():
    """a plan rebound to new f / face arrays and lambdas solves those (same as a fresh single solve)"""
    probs = make(small_descs(12, 21))
    bs = run_batch(probs)
    for p in probs:
        assert rel(p["f"], single(p)) <= SAME
    # new right-hand sides, new face data, new lambdas (same geometry: lambda keeps its zero / non-zero class)
    fresh = []
    for k, p in enumerate(probs):
        d = dict(p["d"])
        if d["lam"] != 0.0:
            d["lam"] = d["lam"] * 0.5 - 0.25
        f = torch.randn_like(p["f0"])
        bcs = [None if b is None else torch.randn_like(b) for b in p["bcs"]]
        fresh.append({"d": d, "f0": f.clone(), "f": f, "bcs": bcs})
    bs.rebind([(q["f"], q["d"]["n"], q["d"]["bc"], q["d"]["lam"], q["bcs"]) for q in fresh])
    bs.solve()
    torch.cuda.synchronize()
    for k, q in enumerate(fresh):
        e = rel(q["f"], single(q))
        assert e <= SAME, (k, q["d"], e)
    # a different geometry is refused
    bad = [(q["f"], q["d"]["n"], q["d"]["bc"], q["d"]["lam"], q["bcs"]) for q in fresh]
    n0 = fresh[0]["d"]["n"]
    bad[0] = (fresh[1]["f"], fresh[1]["d"]["n"], fresh[1]["d"]["bc"], fresh[1]["d"]["lam"], fresh[1]["bcs"])
    if fresh[1]["d"]["n"] != n0 or fresh[1]["d"]["bc"] != fresh[0]["d"]["bc"]:
        with pytest.raises(C.CHZError, match="geometry"):
            bs.rebind(bad)
    bs.close()


@pytest.mark.parametrize("B,n,bc", [(4, (16, 12, 20), (3, 2, 0)), (16, (32, 32, 32), (1, 3, 0))])
def test_solve_batch_cached_matches_single(B, n, bc):
    """helmholtz_solve_batch (cached plan, rebound on the second call) equals single solves"""
    NX, NY, NZ = n
    g = torch.Generator(device=dev())
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    for call in range(3):
        g.manual_seed(100 + call)
        f = torch.randn((B, NZ + 1, NY + 1, NX + 1), generator=g, dtype=torch.float64, device=dev())
        lams = [-1.0 - 0.5 * k - call for k in range(B)]
        bcs = [None] * 6
        for i in range(6):
            if bc[i // 2] in ((3, 4) if i % 2 == 0 else (2, 3)):
                # face 0 / 2 / 4 shared by the batch, 1 / 3 / 5 one per problem
                shape = (sizes[i],) if i % 2 == 0 else (B, sizes[i])
                bcs[i] = torch.randn(shape, generator=g, dtype=torch.float64, device=dev())
        u = f.clone()
        C.helmholtz_solve_batch(u, n, bc, lams, bcs)
        torch.cuda.synchronize()
        for k in range(B):
            fk = [None if b is None else (b[k] if b.dim() == 2 else b) for b in bcs]
            want = f[k].clone()
            C.helmholtz_solve(want, n, bc, lams[k], fk)
            assert rel(u[k], want) <= SAME, (call, k, rel(u[k], want))


def test_adjoint_batch_matches_single():
    """the batched adjoint equals the single-problem adjoint for every problem, gradients of faces and lambda included"""
    n, bc, B = (12, 16, 10), (4, 3, 0), 5
    NX, NY, NZ = n
    g = torch.Generator(device=dev())
    g.manual_seed(31)
    u = torch.randn((B, NZ + 1, NY + 1, NX + 1), generator=g, dtype=torch.float64, device=dev())
    gu = torch.randn((B, NZ + 1, NY + 1, NX + 1), generator=g, dtype=torch.float64, device=dev())
    lams = [-0.5 - k for k in range(B)]
    gf, gfaces, glam = C.helmholtz_adjoint_batch(u, gu, n, bc, lams, faces=True)
    torch.cuda.synchronize()
    for k in range(B):
        g1, f1, l1 = C.helmholtz_adjoint(u[k].contiguous(), gu[k].contiguous(), n, bc, lams[k], faces=True)
        assert rel(gf[k], g1) <= SAME, k
        for i in range(6):
            assert (gfaces[i] is None) == (f1[i] is None)
            if f1[i] is not None:
                assert rel(gfaces[i][k], f1[i]) <= SAME, (k, i)
        assert abs(float(glam[k]) - float(l1)) <= SAME * max(1.0, abs(float(l1))), k


def test_solve_batch_cache_evicts():
    """more geometries than the plan cache holds: every call still solves its own problems"""
    for it in range(2):
        for N in (8, 10, 12, 16, 20, 24):
            n = (N, N, 8)
            f = torch.randn((2, 9, N + 1, N + 1), dtype=torch.float64, device=dev())
            u = f.clone()
            C.helmholtz_solve_batch(u, n, (1, 1, 0), [-1.0, -2.0])
            torch.cuda.synchronize()
            for k in range(2):
                want = f[k].clone()
                C.helmholtz_solve(want, n, (1, 1, 0), [-1.0, -2.0][k])
                assert rel(u[k], want) <= SAME, (it, N, k)


def test_release_then_solve_again():
    """chz_release frees the cached plans and workspaces; later calls rebuild them"""
    n, bc = (16, 12, 8), (1, 3, 0)
    NX, NY, NZ = n
    f = torch.randn((3, NZ + 1, NY + 1, NX + 1), dtype=torch.float64, device=dev())
    face = torch.randn((NZ + 1) * (NX + 1), dtype=torch.float64, device=dev())
    bcs = (None, None, face, face, None, None)
    u1 = f.clone()
    C.helmholtz_solve_batch(u1, n, bc, -1.5, bcs)
    torch.cuda.synchronize()
    C.release()
    u2 = f.clone()
    C.helmholtz_solve_batch(u2, n, bc, -1.5, bcs)
    w = f[1].clone()
    C.helmholtz_solve(w, n, bc, -1.5, [face if b is not None else None for b in bcs])
    torch.cuda.synchronize()
    assert rel(u2, u1) <= SAME and rel(u2[1], w) <= SAME


def test_batch_allocation_failure_reported():
    """A plan whose workspaces exceed the device memory (40 problems of
    1024^3 cells, 350 GB) fails with the allocation status and message, and
    the library keeps working afterwards."""
    import ctypes
    f = torch.zeros(8, dtype=torch.float64, device=dev())   # never read: the workspace allocation comes first
    arr = (C._Problem * 40)()
    for q in arr:
        for i in range(6):
            q.domain[i] = float(i % 2)
            q.faces[i] = None
        for a in range(3):
            q.N[a] = 1024
            q.bc[a] = 1
        q.lam = -1.0
        q.f = f.data_ptr()
    plan = ctypes.c_void_p()
    rc = C.lib().chz_batch_create(arr, 40, 64, 0, ctypes.byref(plan))
    assert rc == -4 and not plan.value
    assert b"allocation failed" in C.lib().chz_last_error()
    p = make(small_descs(3, 14))[0]
    u = single(p)
    torch.cuda.synchronize()
    d = p["d"]
    want = O.helmholtz(p["f_np"], d["n"], d["bc"], d["lam"], p["bcs_np"])
    assert float(np.max(np.abs(u.cpu().numpy() - want)) / np.max(np.abs(want))) <= 1e-12
