"""torch.autograd.Function over the HIP forward and checkpointed adjoint.

    from seisacomod2d_amd.autograd import Seismograms
    seis = Seismograms.apply(vp, rho, modeller_factory, shots, ckpt_k, windows)

forward(vp, rho) runs the per-shot forward through the C ABI and returns the
gathers concatenated as one (sum of nrec, real_nt) tensor; backward(grad)
calls sam2d_gradient with the incoming gradient as the adjoint source and
returns dL/dvp and dL/drho.  Tensors are CPU float32 (x-major (nx, nz)
models, as the reference's model files); the computation itself runs on the
GPU, and there is no CPU fallback.
"""
import numpy as np
import torch

from . import Grid, Modeller, Shot

__all__ = ["Seismograms", "ModellerCache"]


class ModellerCache:
    """Reuses one device context per model value (the model is part of the
    context); a new (vp, rho) pair rebuilds it."""

    def __init__(self, grid: Grid):
        self.grid = grid
        self._key = None
        self._m = None

    def get(self, vp: np.ndarray, rho: np.ndarray) -> Modeller:
        key = (vp.tobytes(), rho.tobytes())
        if self._m is None or self._key != key:
            if self._m is not None:
                self._m.close()
            self._m = Modeller(self.grid, vp, rho)
            self._key = key
        return self._m

    def close(self):
        if self._m is not None:
            self._m.close()
            self._m = None


class Seismograms(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, rho, cache: ModellerCache, shots, ckpt_k=100, windows=None):
        vp_np = vp.detach().cpu().contiguous().numpy().astype(np.float32, copy=False)
        rho_np = rho.detach().cpu().contiguous().numpy().astype(np.float32, copy=False)
        m = cache.get(vp_np, rho_np)
        if windows is None:
            gathers = m.forward(list(shots))
        else:
            gathers = m.forward_windows(list(shots), windows)
        ctx.cache, ctx.shots, ctx.ckpt_k, ctx.windows = cache, shots, ckpt_k, windows
        ctx.models = (vp_np, rho_np)
        out = np.concatenate([g for g in gathers], axis=0)
        return torch.from_numpy(out)

    @staticmethod
    def backward(ctx, grad_out):
        vp_np, rho_np = ctx.models
        m = ctx.cache.get(vp_np, rho_np)
        g = grad_out.detach().cpu().contiguous().numpy().astype(np.float32, copy=False)
        adj, off = [], 0
        for s in ctx.shots:
            adj.append(g[off:off + s.nrec])
            off += s.nrec
        gvp, grho = m.gradient(list(ctx.shots), adj, ckpt_k=ctx.ckpt_k, windows=ctx.windows)
        return torch.from_numpy(gvp), torch.from_numpy(grho), None, None, None, None
