"""Differentiable Helmholtz solve: torch.autograd.Function over the FP64
solve (chz_helmholtz_solve; a batch: chz_batch_solve_cached) and its
hand-written adjoint (chz_helmholtz_adjoint / chz_helmholtz_adjoint_batch).

    u = helmholtz(f, lam, bc, bcs=(None,) * 6, dom=(0, 1) * 3)

f: CUDA float64 node tensor (NZ+1, NY+1, NX+1), or a batch (B, NZ+1, NY+1,
NX+1) of independent problems of one geometry.  lam: float, 0-d tensor or
(B,) tensor.  bcs: six Neumann face tensors (bcl, bcr, bcb, bct, bcf, bce),
None where the codes read none; for a batch each is (size,) (shared by all
problems) or (B, size).  Gradients flow to f, lam and the face tensors:
dL/df and the face gradients are the adjoint solve's, dL/dlam = -<u_bar, u>
with u_bar = A^-T dL/du on the unknown box.
"""
import torch

import cuhelmholtz_amd as C


class HelmholtzSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, lam, bc, dom, *bcs):
        batched = f.dim() == 4
        fb = f.detach().contiguous() if batched else f.detach().contiguous().unsqueeze(0)
        B = fb.shape[0]
        NZ, NY, NX = (int(s) - 1 for s in fb.shape[1:])
        lam_d = lam.detach().to(torch.float64)
        lams = [float(lam_d)] * B if lam_d.dim() == 0 else [float(v) for v in lam_d.reshape(-1).tolist()]
        if len(lams) != B:
            raise ValueError("lam must be a scalar or hold one value per problem")
        u = fb.clone()
        if batched:
            # one batched solve (one graph replay of the library's cached plan for this geometry)
            C.helmholtz_solve_batch(u, (NX, NY, NZ), bc, lams, [None if b is None else b.detach() for b in bcs], dom)
        else:
            faces = [None if b is None else b.detach().contiguous() for b in bcs]
            C.helmholtz_solve(u[0], (NX, NY, NZ), bc, lams[0], faces, dom)
        ctx.geom = ((NX, NY, NZ), tuple(bc), tuple(dom), lams, batched, lam_d.dim())
        ctx.face_shapes = [None if b is None else tuple(b.shape) for b in bcs]
        ctx.save_for_backward(u)
        return u if batched else u[0]

    @staticmethod
    def backward(ctx, gu):
        (u,) = ctx.saved_tensors
        n, bc, dom, lams, batched, lam_dim = ctx.geom
        gub = (gu if batched else gu.unsqueeze(0)).to(torch.float64).contiguous()
        B = u.shape[0]
        want_faces = any(s is not None for s in ctx.face_shapes) and any(ctx.needs_input_grad[4:])
        gf = torch.empty_like(u)
        glam = torch.zeros(B, dtype=torch.float64, device=u.device)
        gfaces = [None if s is None else torch.zeros(s, dtype=torch.float64, device=u.device)
                  for s in ctx.face_shapes]
        if batched:
            gf, fg, gl = C.helmholtz_adjoint_batch(u, gub, n, bc, lams, dom, faces=want_faces,
                                                   want_lambda=ctx.needs_input_grad[1])
            if gl is not None:
                glam = gl
            for i, s in enumerate(ctx.face_shapes):
                if s is None or fg[i] is None:
                    continue
                # per-problem face tensors take their rows, a shared one the sum over the batch
                gfaces[i] = fg[i].reshape(s) if len(s) == 2 else fg[i].sum(0).reshape(s)
        else:
            g, fg, gl = C.helmholtz_adjoint(u[0], gub[0], n, bc, lams[0], dom, faces=want_faces,
                                            want_lambda=ctx.needs_input_grad[1])
            gf[0] = g
            if gl is not None:
                glam[0] = gl
            for i, s in enumerate(ctx.face_shapes):
                if s is not None and fg[i] is not None:
                    gfaces[i] += fg[i].reshape(s)
        g_f = gf if batched else gf[0]
        g_lam = None
        if ctx.needs_input_grad[1]:
            g_lam = glam.sum() if lam_dim == 0 else glam
        return (g_f, g_lam, None, None, *gfaces)


def helmholtz(f, lam, bc, bcs=(None,) * 6, dom=(0.0, 1.0) * 3):
    """Differentiable FP64 solve of laplace u + lam u = f (see the module docstring)."""
    if not torch.is_tensor(lam):
        lam = torch.tensor(float(lam), dtype=torch.float64, device=f.device)
    bcs = list(bcs) + [None] * (6 - len(bcs))
    return HelmholtzSolve.apply(f, lam, tuple(int(b) for b in bc), tuple(float(v) for v in dom), *bcs)
