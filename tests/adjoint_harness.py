"""Test-only torch.autograd wrappers of the library's FP64 projection and
interface updates, with the library's hand-written transposes as their
backward (include/biocfd_hip.h, biocfd_projection_adjoint,
biocfd_coarse_update_adjoint, biocfd_fine_update_bd_adjoint), and of the
whole adjoint workload: the time-averaged body x-force as a function of the
body's kinematic parameters, with the adjoint time steps as its backward
(biocfd_amd.adjoint.AdjointRun).  Fields are
flat float64 CPU tensors in the block layout (i fastest, one ghost layer);
the work runs on the GPU through the C ABI."""
import numpy as np
import torch

from biocfd_amd import hip as H

FIELD = {"p": H.BF["p"], "u": H.BF["u"], "v": H.BF["v"], "w": H.BF["w"]}
COMP_FIELD = ("p", "u", "v", "w")


def _np(t):
    return np.ascontiguousarray(t.detach().numpy(), dtype=np.float64)


def upload(g, name, t):
    H.call("biocfd_field_upload", g, H.BF[name], H.ptr(_np(t)))


def download(g, name, n):
    a = np.zeros(n)
    H.call("biocfd_field_download", g, H.BF[name], H.ptr(a))
    return torch.from_numpy(a)


class Projection(torch.autograd.Function):
    """(ut, vt, wt, p) -> (ut', vt', wt', p') of block g: computeDiv,
    n_iter red-black SOR iterations from pc = 0, correctPressure and
    correctVelocity (biocfd_projection)."""

    @staticmethod
    def forward(ctx, ut, vt, wt, p, g, n_iter, dt):
        n = ut.numel()
        for name, t in (("ut", ut), ("vt", vt), ("wt", wt), ("p", p)):
            upload(g, name, t)
        H.call("biocfd_projection", g, n_iter, dt)
        ctx.g, ctx.n_iter, ctx.dt, ctx.n = g, n_iter, dt, n
        return tuple(download(g, name, n) for name in ("ut", "vt", "wt", "p"))

    @staticmethod
    def backward(ctx, gu, gv, gw, gp):
        n = ctx.n
        outs = [np.zeros(n) for _ in range(3)]
        ins = [_np(x) for x in (gu, gv, gw, gp)]
        H.call("biocfd_projection_adjoint", ctx.g, ctx.n_iter, ctx.dt, *(H.ptr(a) for a in ins),
               *(H.ptr(a) for a in outs))
        return (*(torch.from_numpy(a) for a in outs), gp.clone(), None, None, None)


class Restriction(torch.autograd.Function):
    """(coarse, fine) field of one component -> coarse field after
    coarseUpdate on every interface (biocfd_coarse_update(0), component comp
    of p, u, v, w); interface itf's transpose as backward (one interface)."""

    @staticmethod
    def forward(ctx, coarse, fine, itf, comp, a_blk, b_blk):
        name = COMP_FIELD[comp]
        upload(a_blk, name, coarse)
        upload(b_blk, name, fine)
        H.call("biocfd_coarse_update", 0)
        ctx.itf, ctx.comp, ctx.nf = itf, comp, fine.numel()
        return download(a_blk, name, coarse.numel())

    @staticmethod
    def backward(ctx, gc):
        lc = _np(gc).copy()
        lf = np.zeros(ctx.nf)
        H.call("biocfd_coarse_update_adjoint", ctx.itf, ctx.comp, H.ptr(lc), H.ptr(lf))
        return torch.from_numpy(lc), torch.from_numpy(lf), None, None, None, None


class Prolongation(torch.autograd.Function):
    """(fine, coarse) field of one component -> fine field after
    fineUpdate_bd_mv of interface itf (biocfd_fine_update_bd_one(itf, 0))."""

    @staticmethod
    def forward(ctx, fine, coarse, itf, comp, a_blk, b_blk):
        name = COMP_FIELD[comp]
        upload(a_blk, name, coarse)
        upload(b_blk, name, fine)
        H.call("biocfd_fine_update_bd_one", itf, 0)
        ctx.itf, ctx.comp, ctx.nc = itf, comp, coarse.numel()
        return download(b_blk, name, fine.numel())

    @staticmethod
    def backward(ctx, gf):
        lf = _np(gf).copy()
        lc = np.zeros(ctx.nc)
        H.call("biocfd_fine_update_bd_adjoint", ctx.itf, ctx.comp, H.ptr(lf), H.ptr(lc))
        return torch.from_numpy(lf), torch.from_numpy(lc), None, None, None, None


class BodyForce(torch.autograd.Function):
    """(theta, body nodes) -> J, the mean body x-force over the run's
    averaging window; theta = (heave amplitude, pitch amplitude, phase,
    frequency), nodes (nn, 3) the body-frame marker nodes: forward =
    AdjointRun.forward with checkpoints, backward = the adjoint time steps
    replayed from them (AdjointRun.backward).  A frozen-geometry run has no
    node gradient (None)."""

    @staticmethod
    def forward(ctx, theta, nodes, run):
        nd = nodes.detach().numpy()
        J = run.forward(theta.detach().numpy(), nodes=(nd[:, 0].copy(), nd[:, 1].copy(), nd[:, 2].copy()),
                        checkpoints=True)
        ctx.run = run
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        out = ctx.run.backward()
        if isinstance(out, tuple):
            return gJ * torch.from_numpy(out[0]), gJ * torch.from_numpy(out[1]), None
        return gJ * torch.from_numpy(out), None, None
