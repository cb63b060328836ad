"""Differentiable expectation E = <psi|H|psi> with a hand-written adjoint.

``expectation_value(op, psi, coeffs)`` returns E as a 0-d complex128 tensor
and is differentiable in ``psi`` (complex128, on the operator's device) and in
``coeffs`` (complex128 or float64, one per term of ``op``; when omitted the
operator's own coefficients are used and only ``psi`` is differentiated).

With PyTorch's convention for complex gradients (grad = dL/dRe + i dL/dIm for a
real loss L) and grad_output g for E:

    grad_psi = conj(g) * (H psi) + g * (H^dagger psi) = H_d psi,
               H_d = sum_k d_k P_k with real d_k = 2 Re(conj(g) c_k)
    grad_c   = g * <P_k>               (real part for real coefficients)

so the backward is one application of H_d: one sigma sweep with real
coefficients (no H^dagger psi, no per-term values), and the per-term values
<P_k> are produced by the forward sweep and saved when the coefficients need a
gradient. Forward plus backward therefore costs one forward evaluation and one
sigma sweep.

PyTorch is used only here, at the Python boundary; the arithmetic runs in the
HIP kernels behind the C ABI.
"""

from __future__ import annotations

import weakref
from typing import Any

import torch

from .pauli_sum import DevicePauliSum


def _stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _check_psi(op: DevicePauliSum, psi: torch.Tensor) -> torch.Tensor:
    if not isinstance(psi, torch.Tensor) or not psi.is_cuda:
        raise TypeError("psi must be a device-resident torch tensor")
    if psi.dtype != torch.complex128:
        raise TypeError("psi dtype must be complex128 for the differentiable expectation")
    if psi.dim() != 1:
        raise ValueError("psi must be 1-dimensional")
    if psi.device.index != op.device:
        raise ValueError("psi must be on the operator's device")
    return psi.contiguous()


_OWN_COEFFS: "weakref.WeakKeyDictionary[DevicePauliSum, torch.Tensor]" = weakref.WeakKeyDictionary()


def _own_coeffs(op: DevicePauliSum, device: torch.device) -> torch.Tensor:
    """The operator's own coefficients as a device tensor (copied once per
    operator and kept while the operator lives)."""
    cached = _OWN_COEFFS.get(op)
    if cached is None or cached.device != device:
        cached = torch.from_numpy(op.to_host().coeffs()).to(device)
        _OWN_COEFFS[op] = cached
    return cached


class PauliExpectation(torch.autograd.Function):
    @staticmethod
    def forward(ctx: Any, psi: torch.Tensor, coeffs: torch.Tensor | None, op: DevicePauliSum) -> torch.Tensor:
        psi = _check_psi(op, psi)
        cplx = None
        if coeffs is not None:
            if coeffs.shape != (op.num_terms,):
                raise ValueError("coeffs must have one entry per term")
            cplx = coeffs.detach().to(device=psi.device, dtype=torch.complex128).contiguous()
        want_tv = coeffs is not None and coeffs.requires_grad
        tv = torch.empty(op.num_terms, dtype=torch.complex128, device=psi.device) if want_tv else None
        e = op.expectation_adjoint(
            psi.data_ptr(), psi.numel(), tv.data_ptr() if tv is not None else None, None, None,
            stream=_stream(psi.device), coeffs_ptr=cplx.data_ptr() if cplx is not None else None,
        )
        ctx.op = op
        ctx.coeffs_given = coeffs is not None
        ctx.real_coeffs = coeffs is not None and not coeffs.is_complex()
        empty = torch.empty(0, device=psi.device)
        ctx.save_for_backward(psi, cplx if cplx is not None else empty, tv if tv is not None else empty)
        return torch.tensor(complex(e), dtype=torch.complex128, device=psi.device)

    @staticmethod
    def backward(ctx: Any, grad: torch.Tensor):
        psi, cplx, tv = ctx.saved_tensors
        op: DevicePauliSum = ctx.op
        need_psi = ctx.needs_input_grad[0]
        need_c = ctx.coeffs_given and ctx.needs_input_grad[1]
        g = grad.to(torch.complex128)
        grad_psi = None
        if need_psi:
            c = cplx if ctx.coeffs_given else _own_coeffs(op, psi.device)
            d = (2.0 * (g.conj() * c).real).to(torch.complex128).contiguous()
            grad_psi = torch.empty_like(psi)
            op.expectation_adjoint(psi.data_ptr(), psi.numel(), None, grad_psi.data_ptr(), None,
                                   stream=_stream(psi.device), coeffs_ptr=d.data_ptr(), want_energy=False)
        grad_c = None
        if need_c:
            grad_c = g * tv.conj()
            if ctx.real_coeffs:
                grad_c = grad_c.real
        return grad_psi, grad_c, None


def expectation_value(op: DevicePauliSum, psi: torch.Tensor, coeffs: torch.Tensor | None = None) -> torch.Tensor:
    """E = sum_k c_k <psi|P_k|psi> as a differentiable 0-d complex128 tensor."""
    return PauliExpectation.apply(psi, coeffs, op)
