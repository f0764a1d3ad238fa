"""Differentiable Pk_ideal and Bk_ideal.

``pk_ideal(pspec, field)`` and ``bk_ideal(bspec, field)`` return the ideal estimators of
PolyBin3D's PSpec.Pk_ideal / BSpec.Bk_ideal (pspec.py:754-784, bspec.py:1085-1119) as FP64 device
tensors, concatenated over multipoles as in the reference's output vectors, and are differentiable
with respect to the real float32 field through torch.autograd.  The backward passes are
hand-written adjoints (no counterpart in the reference) built from the same fused FFT, bin and
triplet kernels: the Pk adjoint is one R2C and one C2R with a per-bin multipole weight in its
prologue; the Bk adjoint recomputes the g-maps (nothing is stored between the passes), forms the
adjoint maps with one split-BF16 MFMA contraction and sends them back through the transposed
filters.  The ideal Fisher matrices are data-independent and computed once (as by the reference).
Global sightline (the local-sightline harmonics sums have no adjoint here).
"""
import numpy as np
import torch


def _inv_fish(est, **kw):
    if not hasattr(est, 'inv_fish_ideal'):
        est.compute_fisher_ideal(**kw)
    return torch.from_numpy(np.ascontiguousarray(est.inv_fish_ideal)).to(est.base.device)


class _PkIdeal(torch.autograd.Function):
    @staticmethod
    def forward(ctx, field, ps):
        num = ps._pk_numerator_dev(field, 'ideal')
        ctx.ps = ps
        ctx.save_for_backward(field)
        return _inv_fish(ps) @ num

    @staticmethod
    def backward(ctx, grad):
        (field,) = ctx.saved_tensors
        return ctx.ps._pk_ideal_adjoint(field, grad).reshape(field.shape), None


class _BkIdeal(torch.autograd.Function):
    @staticmethod
    def forward(ctx, field, bs):
        if not hasattr(bs, 'sym_factor'):
            bs._compute_symmetry_factor()
        num = torch.from_numpy(bs.Bk_numerator_ideal(field)).to(bs.base.device)
        ctx.bs = bs
        ctx.save_for_backward(field)
        return _inv_fish(bs) @ num

    @staticmethod
    def backward(ctx, grad):
        (field,) = ctx.saved_tensors
        return ctx.bs._bk_ideal_adjoint(field, grad).reshape(field.shape), None


def pk_ideal(pspec, field):
    """Pk_ideal(field) as a differentiable FP64 tensor [Nl * Nk] (p0, p2, ... concatenated)."""
    if not hasattr(pspec, 'inv_fish_ideal'):
        pspec.compute_fisher_ideal()
    return _PkIdeal.apply(pspec.base.as_real(field), pspec)


def bk_ideal(bspec, field):
    """Bk_ideal(field) as a differentiable FP64 tensor [N_bins] (b0, b2, ... concatenated)."""
    if not hasattr(bspec, 'inv_fish_ideal'):
        bspec.compute_fisher_ideal()
    return _BkIdeal.apply(bspec.base.as_real(field), bspec)
