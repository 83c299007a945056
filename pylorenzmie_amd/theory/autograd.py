"""Holograms as a differentiable torch function of the particle parameters.

``HologramFunction`` runs the FP64 forward kernel (lmz_hologram) and, in
backward, the vector-Jacobian product kernel (lmz_hologram_vjp): the
forward-mode Jacobian of every pixel is formed in registers, contracted
with the upstream gradient and reduced per feature in FP64.  Backward
recomputes the Mie coefficients, their tangents and the per-pixel
recurrences from the saved parameters; nothing but the parameters is kept
between the passes.  The reference takes the same derivatives with
jax.jacfwd (theory/jaxLorenzMie.py:280-286, 354-395).
"""
from __future__ import annotations

import numpy as np

from .. import _lmz
from .batch import HologramBatch


def _torch():
    return _lmz.torch()


class _Buffers:
    """Device tables one HologramBatch needs for a backward pass."""

    def __init__(self, hb: HologramBatch):
        t = _torch()
        b = hb.batch
        dev = hb.params.device
        self.params = t.zeros_like(hb.params)
        self.nmax = t.zeros(hb.nfeat, dtype=t.int32, device=dev)
        self.coef = t.zeros(b.coef_floats(), dtype=t.float32, device=dev)
        self.dcoef = t.zeros(b.dcoef_floats(), dtype=t.float32, device=dev)
        self.tiles = t.zeros(b.ntiles * _lmz.NJAC, dtype=t.float64, device=dev)


def _buffers(hb: HologramBatch) -> _Buffers:
    buf = getattr(hb, '_vjp_buffers', None)
    if buf is None:
        buf = _Buffers(hb)
        hb._vjp_buffers = buf
    return buf


def _hologram_function():
    t = _torch()

    class HologramFunction(t.autograd.Function):
        """I = hologram(params) for the features of a HologramBatch.

        params: (nfeat, 7) float64 tensor on the GPU, rows x_p y_p z_p
        (absolute pixels) a_p n_p k_p alpha.  Returns the packed float32
        holograms (npix_total,); backward gives dL/dparams in float64."""

        @staticmethod
        def forward(ctx, params, hb):
            if params.dim() != 2 or params.shape != (hb.nfeat, _lmz.NJAC):
                raise ValueError(f'params must be ({hb.nfeat}, {_lmz.NJAC})')
            with t.no_grad():
                hb.params[:, :_lmz.NJAC].copy_(params.detach())
            out = hb.compute().clone()
            ctx.hb = hb
            ctx.save_for_backward(params.detach().clone())
            return out

        @staticmethod
        def backward(ctx, grad_out):
            hb = ctx.hb
            (params,) = ctx.saved_tensors
            buf = _buffers(hb)
            lib = _lmz.load_library()
            s = _lmz.stream_handle()
            b = hb.batch
            buf.params.zero_()
            buf.params[:, :_lmz.NJAC].copy_(params)
            g = grad_out.detach().to(t.float32).contiguous()
            grad = t.empty((hb.nfeat, _lmz.NJAC), dtype=t.float64,
                           device=params.device)
            _lmz.check(lib.lmz_mie_ab(b.handle, _lmz.ptr(buf.params), None,
                                      _lmz.ptr(buf.nmax), _lmz.ptr(buf.coef),
                                      _lmz.ptr(buf.dcoef), None,
                                      _lmz.ptr(hb.work), s))
            _lmz.check(lib.lmz_hologram_vjp(
                b.handle, _lmz.ptr(buf.params), _lmz.ptr(buf.nmax),
                _lmz.ptr(buf.coef), _lmz.ptr(buf.dcoef), None, hb.flags,
                _lmz.ptr(g), _lmz.ptr(buf.tiles), _lmz.ptr(grad), s))
            return grad.to(params.dtype), None

    return HologramFunction


_FN = None


def hologram(params, hb: HologramBatch):
    """Differentiable packed holograms of ``hb``'s features at ``params``
    ((nfeat, 7) float64 CUDA tensor, requires_grad as needed)."""
    global _FN
    if _FN is None:
        _FN = _hologram_function()
    return _FN.apply(params, hb)


def loss_and_grad(params: np.ndarray, hb: HologramBatch, data):
    """L = sum (I - d)^2 over all pixels and dL/dparams, for tests and
    first-order fitting: one forward and one VJP launch."""
    t = _torch()
    p = t.tensor(np.asarray(params, float), device=hb.params.device,
                 requires_grad=True)
    I = hologram(p, hb)
    L = ((I.double() - data.double()) ** 2).sum()
    L.backward()
    return float(L.item()), p.grad.detach().cpu().numpy()
