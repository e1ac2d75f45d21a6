"""CPU checks of the test-side torch helpers (fold, residual) and the workload
generators against the oracle, so the full-size GPU checks rest on pinned code."""
import itertools

import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")

from cuhelmholtz_amd import checks as H  # noqa: E402
from cuhelmholtz_amd import workloads as W  # noqa: E402


def faces(rng, n, bc):
    NX, NY, NZ = n
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    return [rng.standard_normal(sizes[i]) if bc[i // 2] in ((3, 4) if i % 2 == 0 else (2, 3)) else None
            for i in range(6)]


def test_torch_fold_and_residual_match_oracle():
    rng = np.random.default_rng(0)
    n = (6, 8, 4)
    dom = (0.0, 2.0, -1.0, 1.0, 0.0, 0.5)
    for bc in itertools.product(range(5), repeat=3):
        f = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
        bcs = faces(rng, n, bc)
        ff_o = O.bc_fold(f, n, bc, bcs, dom)
        ff_t = H.fold(torch.from_numpy(f), n, bc, bcs, dom).numpy()
        assert np.max(np.abs(ff_o - ff_t)) < 1e-12
        u = rng.standard_normal(f.shape)
        ro, r_arr = O.residual(u, ff_o, n, bc, -1.5, dom)
        rt, _ = H.residual(torch.from_numpy(u), torch.from_numpy(ff_o), n, bc, -1.5, dom)
        assert abs(ro - rt) <= 1e-9 * max(ro, 1.0), bc
        # and the oracle's solve has a tiny residual under the torch helper
        us = O.helmholtz(f, n, bc, -1.5, bcs, dom)
        r, scale = H.residual(torch.from_numpy(us), torch.from_numpy(ff_o), n, bc, -1.5, dom)
        assert r / scale < 1e-14


def test_workloads_shapes_and_manufactured_solution():
    f, u, bcs = W.manufactured((16, 16, 16))
    got = O.helmholtz(f, (16, 16, 16), (1, 2, 0), -1.0, bcs)
    assert np.max(np.abs(got - u)) < 2e-2   # O(h^2) discretisation error at h = 1/16
    f3, bcs3 = W.smooth_modes((8, 6, 4))
    assert f3.shape == (5, 7, 9) and bcs3[0].size == 5 * 7 and bcs3[2].size == 5 * 9
    probs = W.batched_problems(64)
    assert len(probs) == 64 and all(p["n"][a] in (32, 48, 64, 96, 128) for p in probs for a in range(3))
    assert all(-10 <= p["lam"] <= 0.5 for p in probs)
