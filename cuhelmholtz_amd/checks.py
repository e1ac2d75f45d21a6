"""Accuracy checks of a solve in torch (FP64): the boundary folding and the
7-point residual of the discrete system the solver inverts (the reference's
cudahelmholtz/_cuda_helmholtz_.inl:70-227 folds; ghost-point Neumann,
periodic wrap, Dirichlet values folded into the rhs).  Used to state the
error of every benchmarked result and as a size-independent test invariant."""
import torch

OFF = {0: (0, -1), 1: (1, -1), 2: (1, 0), 3: (0, 0), 4: (0, -1)}


def windows(n, bc):
    return [(OFF[bc[a]][0], n[a] + OFF[bc[a]][1]) for a in range(3)]


def fold(f, n, bc, bcs, dom=(0.0, 1.0) * 3):
    """f (NZ+1, NY+1, NX+1) -> folded copy (reference order x, y, z)."""
    ff = f.clone()
    win = windows(n, bc)
    for a in range(3):
        if bc[a] == 0:
            continue
        ih = n[a] / (dom[2 * a + 1] - dom[2 * a])
        fa = ff.movedim(2 - a, 0)
        b_ax, c_ax = [x for x in range(3) if x != a]
        cs = slice(win[c_ax][0], win[c_ax][1] + 1)
        bs = slice(win[b_ax][0], win[b_ax][1] + 1)
        first, last = win[a]
        if bc[a] in (1, 2):
            fa[first, cs, bs] -= fa[0, cs, bs] * ih * ih
        else:
            g = torch.as_tensor(bcs[2 * a], dtype=f.dtype, device=f.device).reshape(n[c_ax] + 1, n[b_ax] + 1)
            fa[first, cs, bs] += 2 * g[cs, bs] * ih
        if bc[a] in (1, 4):
            fa[last, cs, bs] -= fa[n[a], cs, bs] * ih * ih
        else:
            g = torch.as_tensor(bcs[2 * a + 1], dtype=f.dtype, device=f.device).reshape(n[c_ax] + 1, n[b_ax] + 1)
            fa[last, cs, bs] -= 2 * g[cs, bs] * ih
    return ff


def residual(u, ffold, n, bc, lam, dom=(0.0, 1.0) * 3):
    """(max |ffold - A u| over the unknown box, max|A u|-scale normaliser)."""
    win = windows(n, bc)
    sl = tuple(slice(win[2 - d][0], win[2 - d][1] + 1) for d in range(3))
    U = u[sl]
    Au = lam * U
    for a in range(3):
        d = 2 - a
        ih2 = (n[a] / (dom[2 * a + 1] - dom[2 * a])) ** 2
        m = U.shape[d]
        prev = torch.roll(U, 1, d)
        nxt = torch.roll(U, -1, d)
        first = [slice(None)] * 3
        last = [slice(None)] * 3
        first[d] = 0
        last[d] = m - 1
        first, last = tuple(first), tuple(last)
        if bc[a] != 0:
            if bc[a] in (3, 4):
                prev[first] = U.narrow(d, 1, 1).squeeze(d) if m > 1 else 0
            else:
                prev[first] = 0
            if bc[a] in (2, 3):
                nxt[last] = U.narrow(d, m - 2, 1).squeeze(d) if m > 1 else 0
            else:
                nxt[last] = 0
        Au = Au + (prev - 2 * U + nxt) * ih2
    r = ffold[sl] - Au
    anorm = abs(lam) + 4 * sum((n[a] / (dom[2 * a + 1] - dom[2 * a])) ** 2 for a in range(3))
    scale = anorm * U.abs().max().item() + ffold[sl].abs().max().item()
    return r.abs().max().item(), scale


def relative_residual(u, f_orig, n, bc, lam, bcs=(None,) * 6, dom=(0.0, 1.0) * 3):
    """|f - A u|_inf / (|A| |u|_inf + |f|_inf) of solution u for rhs f_orig."""
    r, scale = residual(u, fold(f_orig, n, bc, bcs, dom), n, bc, lam, dom)
    return r / scale
