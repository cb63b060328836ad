"""Hand-written adjoint behind torch.autograd: gradients against the oracle's
H psi / per-term values (1e-12 relative) and against central finite differences
(h = 1e-6; truncation plus rounding ~ eps |E| / h, tolerance 1e-7 absolute)."""

import numpy as np
import pytest

from conftest import packed
from wolfgang_amd.workloads import chemistry_masks, random_pauli_masks, random_statevector

pytestmark = pytest.mark.gpu
H_FD = 1e-6


def loss_fn(e):
    # A real loss touching both parts of E.
    return e.real + 0.3 * e.imag


@pytest.mark.parametrize("gen,n", [("random", 8), ("chemistry", 8), ("random", 18), ("chemistry", 19)])
def test_gradients_match_oracle_and_finite_differences(gpu, orc, gen, n):
    # n = 18, 19: tiled states (owner-sweep forward, one tiled sigma sweep
    # with real coefficients in the backward).
    import torch

    from wolfgang_amd.autograd import expectation_value

    terms = 60 if n < 18 else 400
    op = random_pauli_masks(n, terms, 0.4, 21) if gen == "random" else chemistry_masks(n, terms, 22)
    dop = op.to_device()
    psi_np = random_statevector(n, 23)
    psi = torch.from_numpy(psi_np).to("cuda:0").requires_grad_(True)
    if gen == "random":
        coeffs = torch.from_numpy(op.coeffs()).to("cuda:0").requires_grad_(True)
    else:
        coeffs = torch.from_numpy(op.coeffs().real.copy()).to("cuda:0").requires_grad_(True)
    e = expectation_value(dop, psi, coeffs)
    assert dop.expectation_plan_info()["tiled"] == (n >= 18)
    want_e = orc.expectation_compensated(*packed(op), psi_np, threads=8)
    scale = float(np.abs(op.coeffs()).sum())
    assert abs(complex(e.item()) - want_e) <= 1e-12 * scale
    loss = loss_fn(e)
    loss.backward()
    # Analytic reference from the oracle: g = dL/dRe E + i dL/dIm E = 1 + 0.3i.
    g = 1.0 + 0.3j
    sigma = orc.apply(*packed(op), psi_np, threads=8)
    sigma_dag = orc.apply(n, op.x_words(), op.z_words(), op.coeffs().conj(), psi_np, threads=8)
    want_gpsi = np.conj(g) * sigma + g * sigma_dag
    np.testing.assert_allclose(psi.grad.cpu().numpy(), want_gpsi, rtol=0, atol=1e-12 * scale)
    tv = orc.term_values(n, op.x_words(), op.z_words(), psi_np, threads=8)
    want_gc = g * np.conj(tv)
    got_gc = coeffs.grad.cpu().numpy()
    if gen == "random":
        np.testing.assert_allclose(got_gc, want_gc, rtol=0, atol=1e-12)
    else:
        np.testing.assert_allclose(got_gc, want_gc.real, rtol=0, atol=1e-12)

    # Central finite differences on a few coefficients and amplitudes.
    def value(p, c):
        with torch.no_grad():
            return loss_fn(expectation_value(dop, p, c)).item()

    rng = np.random.default_rng(5)
    with torch.no_grad():
        for k in rng.choice(op.num_terms, size=4, replace=False):
            d = torch.zeros_like(coeffs)
            d[k] = H_FD
            fd = (value(psi, coeffs + d) - value(psi, coeffs - d)) / (2 * H_FD)
            assert fd == pytest.approx(float(np.real(got_gc[k])), abs=1e-7)
        for j in rng.choice(psi.numel(), size=4, replace=False):
            for unit, part in ((1.0, "re"), (1j, "im")):
                d = torch.zeros_like(psi)
                d[j] = unit * H_FD
                fd = (value(psi + d, coeffs) - value(psi - d, coeffs)) / (2 * H_FD)
                grad = psi.grad[j].item()
                want = grad.real if part == "re" else grad.imag
                assert fd == pytest.approx(want, abs=1e-7)


@pytest.mark.parametrize("n", [6, 18])
def test_gradient_without_coefficient_tensor(gpu, orc, n):
    import torch

    from wolfgang_amd.autograd import expectation_value

    op = chemistry_masks(n, 40, 31)
    psi_np = random_statevector(n, 32)
    psi = torch.from_numpy(psi_np).to("cuda:0").requires_grad_(True)
    e = expectation_value(op.to_device(), psi)
    e.real.backward()
    sigma = orc.apply(*packed(op), psi_np, threads=8)
    # Hermitian H, real loss Re E: grad = 2 H psi.
    np.testing.assert_allclose(psi.grad.cpu().numpy(), 2 * sigma, rtol=0, atol=1e-12 * 40)
