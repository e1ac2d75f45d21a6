"""GPU checks of the autotuner's configurations: the tuned launch shapes (the
fixed-shape kernels compiled from the tuning table) and the built-in choice
(generic kernels where no shape is compiled) both match the oracle, single
and batched, FP64 (1e-12) and FP32 + 1 refinement (1e-10 against FP64); the
tuner itself, run on a restricted candidate set, pins a configuration whose
solve matches the oracle."""
import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not torch.cuda.is_available(), reason="no GPU")]

import cuhelmholtz_amd as C  # noqa: E402
from cuhelmholtz_amd import autotune as A  # noqa: E402
from cuhelmholtz_amd import kernelgen as G  # noqa: E402

TOL = 1e-12


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def faces(rng, n, bc):
    NX, NY, NZ = n
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    return [rng.standard_normal(sizes[i]) if bc[i // 2] in ((3, 4) if i % 2 == 0 else (2, 3)) else None
            for i in range(6)]


@pytest.fixture
def tuned_state():
    """restores the shipped tuning table after the test"""
    yield
    C.reset_axis_configs()
    C.load_tuning()


def _kernel_kinds(n, prec, bc):
    """per role: is the current configuration of (N, prec, role) compiled?"""
    have = {(k[1], k[2], k[3], k[4], k[5], k[6]) for k in C.compiled_kernels() if k[0] in (G.F_LINE, G.F_FIXED)}
    out = []
    zrole = C.ROLE_ZMID if bc[2] == 0 else C.ROLE_ZMIDN
    for role, N in ((C.ROLE_XFWD, n[0]), (C.ROLE_COLS, n[1]), (zrole, n[2]), (C.ROLE_ROWS, n[0])):
        cfg = C.axis_config(N, prec, role=role)
        out.append(all((prec, op, N, cfg["ept"], cfg["tpl"], cfg["lines"]) in have for op in G.role_ops(role, prec)))
    return out


CASES = [((64, 48, 96), (3, 4, 0), -4.0), ((96, 64, 48), (1, 2, 3), 0.5), ((128, 32, 64), (4, 0, 2), -7.5)]


@pytest.mark.parametrize("n,bc,lam", CASES)
def test_tuned_and_builtin_match_oracle(tuned_state, n, bc, lam):
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(sum(n) + bc[0])
    f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
    bcs = faces(rng, n, bc)
    want = O.helmholtz(f, n, bc, lam, bcs)
    bct = [None if b is None else torch.from_numpy(b).to(dev) for b in bcs]
    outs = {}
    for label in ("tuned", "builtin"):
        if label == "tuned":
            C.reset_axis_configs()
            C.load_tuning()
            assert all(_kernel_kinds(n, 64, bc)), "tuned configurations run compiled kernels"
        else:
            C.reset_axis_configs()
            assert not all(_kernel_kinds(n, 64, bc)), "the built-in choice differs somewhere"
        u = torch.from_numpy(f).to(dev)
        C.helmholtz_solve(u, n, bc, lam, bct)
        m = torch.from_numpy(f).to(dev)
        C.helmholtz_solve_mixed(m, n, bc, lam, bct, refine=1)
        torch.cuda.synchronize()
        u, m = u.cpu().numpy(), m.cpu().numpy()
        assert rel(u, want) < TOL, (label, rel(u, want))
        assert rel(m, u) < 1e-10, (label, rel(m, u))
        outs[label] = u
    assert rel(outs["tuned"], outs["builtin"]) < TOL


def test_tuned_and_builtin_batched_match_oracle(tuned_state):
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(11)
    probs, wants, fs = [], [], []
    for n, bc, lam in CASES:
        f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
        bcs = faces(rng, n, bc)
        wants.append(O.helmholtz(f, n, bc, lam, bcs))
        fs.append(f)
        probs.append([None, n, bc, lam, [None if b is None else torch.from_numpy(b).to(dev) for b in bcs]])
    for setup in (lambda: (C.reset_axis_configs(), C.load_tuning()), C.reset_axis_configs):
        setup()
        for prec, refine in ((64, 0), (32, 1)):
            for p, f in zip(probs, fs):
                p[0] = torch.from_numpy(f).to(dev)
            bs = C.BatchSolver([tuple(p) for p in probs], prec, refine)
            bs.solve()
            torch.cuda.synchronize()
            for p, w in zip(probs, wants):
                got = p[0].cpu().numpy()
                assert rel(got, w) < (TOL if prec == 64 else 1e-10), (prec, p[1], rel(got, w))
            bs.close()


def test_transform_tuned_and_builtin_match_oracle(tuned_state):
    dev = torch.device("cuda:0")
    n = (96, 64, 128)
    kinds = (C.DCT_2, C.DST_1, C.DFT_R2C)
    rng = np.random.default_rng(5)
    v = np.zeros((n[2] + 2, n[1] + 2, n[0] + 2))
    wi = [C.in_window(k, N) for k, N in zip(kinds, n)]
    v[wi[2][0]:wi[2][1] + 1, wi[1][0]:wi[1][1] + 1, wi[0][0]:wi[0][1] + 1] = rng.uniform(
        -1, 1, (wi[2][1] - wi[2][0] + 1, wi[1][1] - wi[1][0] + 1, wi[0][1] - wi[0][0] + 1))
    ref = O.transform(v, *kinds)
    lo = [C.out_window(k, N) for k, N in zip(kinds, n)]
    sl = (slice(lo[2][0], lo[2][1] + 1), slice(lo[1][0], lo[1][1] + 1), slice(lo[0][0], lo[0][1] + 1))
    for setup in (lambda: (C.reset_axis_configs(), C.load_tuning()), C.reset_axis_configs):
        setup()
        x = torch.from_numpy(v).to(dev)
        y = torch.empty_like(x)
        C.do_transform(x, y, kinds)
        torch.cuda.synchronize()
        assert rel(y.cpu().numpy()[sl], ref[sl]) < TOL


def test_tuner_pins_a_checked_configuration(tuned_state):
    # the tuner on two roles of a 64^3 solve, over the compiled shapes; the
    # configuration it pins gives a solve that matches the oracle
    n, bc, lam = (64, 64, 64), (3, 4, 0), -4.0
    for role in (C.ROLE_ZMID, C.ROLE_XFWD):
        assert A.compiled_shapes(64, 64, role), "compiled shapes for the 64-cell roles"
    rows = A.tune_solve(n, 64, bc, lam, log=lambda *a: None, compiled_only=True,
                        roles=(C.ROLE_ZMID, C.ROLE_XFWD), reps=2)
    assert [r[0] for r in rows] == [C.ROLE_ZMID, C.ROLE_XFWD]
    for role, N, prec, ept, tpl, lines, ms, res in rows:
        cfg = C.axis_config(N, prec, role=role)
        assert (cfg["ept"], cfg["tpl"], cfg["lines"]) == (ept, tpl, lines)
        assert ms > 0 and res[0][0] == ms
    from cuhelmholtz_amd import workloads as W
    f, bcs = W.smooth_modes(n)
    want = O.helmholtz(f, n, bc, lam, bcs)
    dev = torch.device("cuda:0")
    u = torch.from_numpy(f).to(dev)
    C.helmholtz_solve(u, n, bc, lam, [None if b is None else torch.from_numpy(b).to(dev) for b in bcs])
    torch.cuda.synchronize()
    assert rel(u.cpu().numpy(), want) < TOL
