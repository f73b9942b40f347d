"""torch.autograd.Function over the HIP forward and checkpointed adjoint.

    from seisacomod2d_amd.autograd import Seismograms
    seis = Seismograms.apply(vp, rho, modeller_cache, shots, ckpt_k, windows)

forward(vp, rho) models the shots with the gradient's forward sweep
(sam2d_forward_checkpointed: the seismograms, and the wavefield checkpoints
every ckpt_k steps kept on the device) and returns the gathers concatenated as
one (sum of nrec, real_nt) tensor; backward(grad) runs the backward pass from
those checkpoints (sam2d_gradient_checkpointed) with the incoming gradient as
the adjoint source and returns dL/dvp and dL/drho, so a misfit evaluation and
its gradient cost one forward sweep, not two.  Tensors are CPU float32
(x-major (nx, nz) models, as the reference's model files); the computation
itself runs on the GPU, and there is no CPU fallback.
"""
import numpy as np
import torch

from . import Grid, Modeller

__all__ = ["Seismograms", "ModellerCache"]


class ModellerCache:
    """Reuses one device context while the model is unchanged (the model is
    part of the context); a new (vp, rho) pair rebuilds it."""

    def __init__(self, grid: Grid):
        self.grid = grid
        self._models = None
        self._m = None

    def get(self, vp: np.ndarray, rho: np.ndarray) -> Modeller:
        same = (self._models is not None and np.array_equal(self._models[0], vp)
                and np.array_equal(self._models[1], rho))
        if self._m is None or not same:
            if self._m is not None:
                self._m.close()
            self._m = Modeller(self.grid, vp, rho)
            self._models = (vp.copy(), rho.copy())
        return self._m

    def close(self):
        if self._m is not None:
            self._m.close()
            self._m = None
            self._models = None


class Seismograms(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, rho, cache: ModellerCache, shots, ckpt_k=100, windows=None):
        vp_np = vp.detach().cpu().contiguous().numpy().astype(np.float32, copy=False)
        rho_np = rho.detach().cpu().contiguous().numpy().astype(np.float32, copy=False)
        m = cache.get(vp_np, rho_np)
        gathers = m.forward_checkpointed(list(shots), ckpt_k=ckpt_k, windows=windows)
        ctx.cache, ctx.shots, ctx.ckpt_k, ctx.windows = cache, shots, ckpt_k, windows
        ctx.modeller, ctx.checkpoint_id = m, m.checkpoint_id
        ctx.models = (vp_np, rho_np)
        out = np.concatenate([g for g in gathers], axis=0)
        return torch.from_numpy(out)

    @staticmethod
    def backward(ctx, grad_out):
        g = grad_out.detach().cpu().contiguous().numpy().astype(np.float32, copy=False)
        adj, off = [], 0
        for s in ctx.shots:
            adj.append(g[off:off + s.nrec])
            off += s.nrec
        m = ctx.cache.get(*ctx.models)
        gvp = grho = None
        if m is ctx.modeller and m.checkpoint_id == ctx.checkpoint_id:
            gvp, grho = m.gradient_checkpointed(adj)
        if gvp is None:
            gvp, grho = m.gradient(list(ctx.shots), adj, ckpt_k=ctx.ckpt_k, windows=ctx.windows)
        return torch.from_numpy(gvp), torch.from_numpy(grho), None, None, None, None
