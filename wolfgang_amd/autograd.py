"""Differentiable expectation E = <psi|H|psi> with a hand-written adjoint.

``expectation_value(op, psi, coeffs)`` returns E as a 0-d complex128 tensor
and is differentiable in ``psi`` (complex128, on the operator's device) and in
``coeffs`` (complex128 or float64, one per term of ``op``; when omitted the
operator's own coefficients are used and only ``psi`` is differentiated).

Backward runs the library's adjoint entry point once: it returns the per-term
values <P_k> (dE/dc_k) and sigma = H psi (dE/dpsi*), plus H^dagger psi when the
coefficients are complex. With PyTorch's convention for complex gradients
(grad = dL/dRe + i dL/dIm for a real loss L) and grad_output g for E:

    grad_psi = conj(g) * (H psi) + g * (H^dagger psi)
    grad_c   = g * conj(<P>)          (real part for real coefficients)

PyTorch is used only here, at the Python boundary; the arithmetic runs in the
HIP kernels behind the C ABI.
"""

from __future__ import annotations

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


class PauliExpectation(torch.autograd.Function):
    @staticmethod
    def forward(ctx: Any, psi: torch.Tensor, coeffs: torch.Tensor | None, op: DevicePauliSum) -> torch.Tensor:
        psi = _check_psi(op, psi)
        cplx = None
        if coeffs is not None:
            if coeffs.shape != (op.num_terms,):
                raise ValueError("coeffs must have one entry per term")
            cplx = coeffs.detach().to(device=psi.device, dtype=torch.complex128).contiguous()
        e = op.expectation_adjoint(
            psi.data_ptr(), psi.numel(), None, None, None, stream=_stream(psi.device),
            coeffs_ptr=cplx.data_ptr() if cplx is not None else None,
        )
        ctx.op = op
        ctx.coeffs_given = coeffs is not None
        ctx.real_coeffs = coeffs is not None and not coeffs.is_complex()
        ctx.save_for_backward(psi, cplx if cplx is not None else torch.empty(0, device=psi.device))
        return torch.tensor(complex(e), dtype=torch.complex128, device=psi.device)

    @staticmethod
    def backward(ctx: Any, grad: torch.Tensor):
        psi, cplx = ctx.saved_tensors
        op: DevicePauliSum = ctx.op
        need_psi = ctx.needs_input_grad[0]
        need_c = ctx.coeffs_given and ctx.needs_input_grad[1]
        if not (need_psi or need_c):
            return None, None, None
        sigma = torch.empty_like(psi) if need_psi else None
        # H^dagger psi differs from H psi only for complex coefficients.
        if ctx.coeffs_given:
            hermitian = ctx.real_coeffs
        else:
            hermitian = False
        sigma_dag = torch.empty_like(psi) if (need_psi and not hermitian) else None
        tv = torch.empty(op.num_terms, dtype=torch.complex128, device=psi.device) if need_c else None
        op.expectation_adjoint(
            psi.data_ptr(), psi.numel(),
            tv.data_ptr() if tv is not None else None,
            sigma.data_ptr() if sigma is not None else None,
            sigma_dag.data_ptr() if sigma_dag is not None else None,
            stream=_stream(psi.device),
            coeffs_ptr=cplx.data_ptr() if ctx.coeffs_given else None,
        )
        g = grad.to(torch.complex128)
        grad_psi = None
        if need_psi:
            if sigma_dag is None:
                grad_psi = (g.conj() + g) * sigma
            else:
                grad_psi = g.conj() * sigma + g * sigma_dag
        grad_c = None
        if need_c:
            grad_c = g * tv.conj()
            if ctx.real_coeffs:
                grad_c = grad_c.real
        return grad_psi, grad_c, None


def expectation_value(op: DevicePauliSum, psi: torch.Tensor, coeffs: torch.Tensor | None = None) -> torch.Tensor:
    """E = sum_k c_k <psi|P_k|psi> as a differentiable 0-d complex128 tensor."""
    return PauliExpectation.apply(psi, coeffs, op)
