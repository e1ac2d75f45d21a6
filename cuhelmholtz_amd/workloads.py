"""Synthetic inputs of the benchmark workloads (numpy, seeded).

manufactured_64: u = sin(pi x) sin(pi y/2) cos(2 pi z) on [0,1]^3, codes
    (1,2,0), lambda = -1; f = Delta u + lambda u at the nodes with the
    Dirichlet values in the boundary nodes, like the reference's driver
    (cudahelmholtz/testgpu.cu:57-122).
smooth_modes: the mixed-precision workload: f = sum of 8 products of
    cos(pi p x) cos(pi q y) cos(2 pi r z) with p, q, r ~ U{1..16},
    amplitudes N(0,1) (numpy default_rng(1), drawn p, q, r, a in that order)
    plus 1e-3 * U[-1,1] noise per node; codes (3,4,0) with zero Neumann data.
batched_problems: per-problem sizes from {32,48,64,96,128}, codes 0-4,
    lambda ~ U[-10, 0.5] away from singular / near-resonant values,
    f ~ N(0,1) (default_rng(2)).
"""
import numpy as np


def node_grid(n, dom=(0.0, 1.0) * 3):
    NX, NY, NZ = n
    x = dom[0] + (dom[1] - dom[0]) * np.arange(NX + 1) / NX
    y = dom[2] + (dom[3] - dom[2]) * np.arange(NY + 1) / NY
    z = dom[4] + (dom[5] - dom[4]) * np.arange(NZ + 1) / NZ
    return x, y, z


def manufactured(n, lam=-1.0):
    """codes (1,2,0): returns f (nodes, (NZ+1,NY+1,NX+1)), exact u, bcs tuple."""
    x, y, z = node_grid(n)
    NX, NY, NZ = n
    sx, sy, cz = np.sin(np.pi * x), np.sin(np.pi * y / 2), np.cos(2 * np.pi * z)
    u = cz[:, None, None] * sy[None, :, None] * sx[None, None, :]
    f = (-np.pi ** 2 - np.pi ** 2 / 4 - 4 * np.pi ** 2 + lam) * u
    f[:, :, 0] = u[:, :, 0]
    f[:, :, NX] = u[:, :, NX]
    f[:, 0, :] = u[:, 0, :]
    # y = 1 Neumann: du/dy = (pi/2) sin(pi x) cos(pi/2) cos(2 pi z), face [k*(NX+1)+i]
    bct = ((np.pi / 2) * np.cos(np.pi / 2) * cz[:, None] * sx[None, :]).ravel()
    return f, u, (None, None, None, bct, None, None)


def smooth_modes(n, seed=1, noise=1e-3):
    """Mixed-precision workload rhs on the nodes of [0,1]^3 (codes (3,4,0))."""
    rng = np.random.default_rng(seed)
    p = rng.integers(1, 17, 8)
    q = rng.integers(1, 17, 8)
    r = rng.integers(1, 17, 8)
    a = rng.standard_normal(8)
    x, y, z = node_grid(n)
    NX, NY, NZ = n
    f = np.zeros((NZ + 1, NY + 1, NX + 1))
    for m in range(8):
        f += a[m] * np.cos(2 * np.pi * r[m] * z)[:, None, None] * np.cos(np.pi * q[m] * y)[None, :, None] \
            * np.cos(np.pi * p[m] * x)[None, None, :]
    f += noise * rng.uniform(-1.0, 1.0, f.shape)
    bcl = np.zeros((NZ + 1) * (NY + 1))
    bcr = np.zeros((NZ + 1) * (NY + 1))
    bcb = np.zeros((NZ + 1) * (NX + 1))
    return f, (bcl, bcr, bcb, None, None, None)


def lowest_eigs(bc, N, h):
    """Discrete eigenvalues (2cos(theta_m)-2)/h^2 of one axis (all modes)."""
    m = np.arange(N + 2)
    if bc == 0:
        th = 2 * np.pi * (m // 2) / N
    elif bc == 2:
        th = np.pi * m / N - 0.5 * np.pi / N
    elif bc == 4:
        th = np.pi * m / N + 0.5 * np.pi / N
    else:
        th = np.pi * m / N
    return np.unique((2 * np.cos(th) - 2) / h ** 2)


def batched_problems(count=2048, seed=2, sizes=(32, 48, 64, 96, 128), delta=1e-3):
    """Descriptors of the batched workload: list of dicts (n, bc, lam) and a seed per problem."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = tuple(int(v) for v in rng.choice(sizes, 3))
        bc = tuple(int(v) for v in rng.integers(0, 5, 3))
        while True:
            lam = float(rng.uniform(-10.0, 0.5))
            if all(b in (0, 3) for b in bc) and abs(lam) < 1e-6:
                continue
            ex, ey, ez = (lowest_eigs(bc[a], n[a], 1.0 / n[a]) for a in range(3))
            s = (ex[:, None, None] + ey[None, :, None] + ez[None, None, :]).ravel()
            if np.min(np.abs(s + lam)) < delta:
                continue
            break
        out.append({"n": n, "bc": bc, "lam": lam, "seed": int(rng.integers(0, 2 ** 31))})
    return out


def problem_arrays(desc):
    """f ~ N(0,1) on the nodes and N(0,1) Neumann face data (where the codes
    need it) of one batched problem, from its own seed."""
    NX, NY, NZ = desc["n"]
    bc = desc["bc"]
    rng = np.random.default_rng(desc["seed"])
    f = rng.standard_normal((NZ + 1, NY + 1, NX + 1))
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    bcs = [rng.standard_normal(sizes[i]) if bc[i // 2] in ((3, 4) if i % 2 == 0 else (2, 3)) else None
           for i in range(6)]
    return f, bcs
