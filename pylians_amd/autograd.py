"""Differentiable P0(k) of a particle set: a torch.autograd.Function around the hand-written
adjoint (pyl_pk_adjoint).  Builder-defined; the reference has no gradients.

    P0 = pk0(pos, BoxSize, dims, MAS="CIC", nbins=64)     # (nbins,) float64, differentiable
    loss = (w * P0).sum(); loss.backward()                 # pos.grad: (n, 3) float32

Forward: fused deposit -> delta = rho/rhobar - 1 -> r2c -> window deconvolution -> binning
(float32 grid and FFT, float64 shell sums).  Backward: the same forward recomputed on the GPU, the
gradient spectrum H_k = (dL/dP0_b) (L/N^2)^3 / Nmodes_b c_k^2 delta_k, an unnormalised c2r and the
stencil-derivative gather at the particles.
"""
import torch

from . import _lib
from .Pk_library import pk_layout


def _kmax(dims):
    return pk_layout(dims)[2]


class _P0(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, BoxSize, dims, scheme, nbins):
        p = pos.detach().to(torch.float32).contiguous()
        _, _, kmax, nout = pk_layout(dims)
        out = torch.empty(nout, dtype=torch.float64, device=p.device)
        _lib.check(_lib.lib().pyl_pk_particles(_lib.ptr(p), None, int(p.shape[0]), int(dims), float(BoxSize),
                                               int(scheme), 0, _lib.F32, 0, _lib.ptr(out), _lib.stream()))
        ctx.save_for_backward(p)
        ctx.cfg = (float(BoxSize), int(dims), int(scheme), int(nbins))
        return out[kmax:kmax + nbins].clone()

    @staticmethod
    def backward(ctx, gP0):
        (p,) = ctx.saved_tensors
        BoxSize, dims, scheme, nbins = ctx.cfg
        w = gP0.detach().to(torch.float64).contiguous()
        grad = torch.empty_like(p)
        _, _, _, nout = pk_layout(dims)
        out = torch.empty(nout, dtype=torch.float64, device=p.device)
        _lib.check(_lib.lib().pyl_pk_adjoint(_lib.ptr(p), int(p.shape[0]), dims, BoxSize, scheme, _lib.ptr(w),
                                             nbins, _lib.ptr(grad), _lib.ptr(out), _lib.stream()))
        return grad, None, None, None, None


def pk0(pos, BoxSize, dims, MAS="CIC", nbins=None):
    """Monopole P0 of the first `nbins` bins (|k| in [b+1, b+2) kF), differentiable w.r.t. pos."""
    _lib.require_gpu()
    if MAS not in _lib.SCHEMES:
        raise ValueError("unknown MAS %r" % (MAS,))
    kmax = _kmax(dims)
    nbins = kmax if nbins is None else int(nbins)
    if not 0 < nbins <= kmax:
        raise ValueError("nbins must be in [1, %d]" % kmax)
    if not (isinstance(pos, torch.Tensor) and pos.is_cuda and pos.ndim == 2 and pos.shape[1] == 3):
        raise ValueError("pos must be a (n, 3) tensor on the GPU")
    return _P0.apply(pos, BoxSize, dims, _lib.SCHEMES[MAS], nbins)


def grad_loss(pos, BoxSize, dims, w, MAS="CIC"):
    """dL/dpos for L = sum_b w_b P0_b without autograd bookkeeping; returns (grad, P0[:len(w)])."""
    p = _lib.to_dev(pos, torch.float32)
    wt = _lib.to_dev(w, torch.float64)
    _, _, kmax, nout = pk_layout(dims)
    grad = torch.empty_like(p)
    out = torch.empty(nout, dtype=torch.float64, device=p.device)
    _lib.check(_lib.lib().pyl_pk_adjoint(_lib.ptr(p), int(p.shape[0]), int(dims), float(BoxSize),
                                         _lib.SCHEMES[MAS], _lib.ptr(wt), int(wt.numel()), _lib.ptr(grad),
                                         _lib.ptr(out), _lib.stream()))
    return grad, out[kmax:kmax + int(wt.numel())]
