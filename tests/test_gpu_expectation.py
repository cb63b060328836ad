"""GPU parity for statevector expectation and its adjoint (FP64; 1e-12 relative).

Tolerance: |E_gpu - E_ref| <= 1e-12 * scale with scale = sum_k |c_k| * ||psi||^2,
the bound on |E| from |<P_k>| <= ||psi||^2; E_ref is the oracle's compensated
(double-double) sum, so the tolerance measures the device error alone. Per-term
values are checked against the oracle's compensated per-term sums with the same
relative bound (scale ||psi||^2)."""

import json
import pathlib

import numpy as np
import pytest

import kat_datasets as kd
from conftest import packed
from wolfgang_amd import PauliSum
from wolfgang_amd.workloads import chemistry_masks, random_pauli_masks, random_statevector

pytestmark = pytest.mark.gpu
KATS = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_kats.json").read_text())
RTOL = 1e-12


def scale(op, psi):
    return float(np.abs(op.coeffs()).sum() * np.vdot(psi, psi).real)


@pytest.mark.parametrize("n", [0, 1, 2, 7, 12, 16])
@pytest.mark.parametrize("gen", ["random", "chemistry", "diagonal"])
def test_expectation_matches_oracle(gpu, orc, n, gen):
    if gen == "chemistry" and n < 2:
        pytest.skip("chemistry generator needs two qubits")
    if gen == "random":
        op = random_pauli_masks(n, 300, 0.4, 17 * n + 1)
    elif gen == "chemistry":
        op = chemistry_masks(n, 300, 17 * n + 2)
    else:
        r = random_pauli_masks(n, 120, 0.5, 17 * n + 3)
        op = PauliSum.from_packed(n, np.zeros_like(r.x_words()), r.z_words() | r.x_words(), r.coeffs())
    psi = random_statevector(n, 99 + n)
    want = orc.expectation_compensated(*packed(op), psi)
    got = op.to_device().expectation_statevector(psi)
    assert abs(got - want) <= RTOL * scale(op, psi)
    # Also the plain reference-order oracle (sequential FP64 sum).
    assert abs(got - orc.expectation(*packed(op), psi)) <= 4 * RTOL * scale(op, psi)
    # Deterministic: identical bits on repeat.
    dop = op.to_device()
    assert dop.expectation_statevector(psi) == dop.expectation_statevector(psi)


def test_expectation_complex64_and_device_tensor(gpu, orc):
    import torch

    n = 10
    op = random_pauli_masks(n, 256, 0.4, 5)
    psi = random_statevector(n, 6)
    dop = op.to_device()
    want64 = orc.expectation(*packed(op), psi.astype(np.complex64))
    got64 = dop.expectation_statevector(psi.astype(np.complex64))
    assert abs(got64 - want64) <= RTOL * scale(op, psi)
    t = torch.from_numpy(psi).to("cuda:0")
    assert abs(dop.expectation_statevector(t) - dop.expectation_statevector(psi)) == 0
    t64 = torch.from_numpy(psi.astype(np.complex64)).to("cuda:0")
    assert dop.expectation_statevector(t64) == got64

    class DLPackOnly:
        # A producer that exposes only the DLPack protocol (kDLROCM device memory).
        def __init__(self, tensor):
            self._t = tensor

        def __dlpack__(self, stream=None):
            return self._t.__dlpack__()

        def __dlpack_device__(self):
            return self._t.__dlpack_device__()

    assert dop.expectation_statevector(DLPackOnly(t)) == dop.expectation_statevector(t)
    assert int(t.__dlpack_device__()[0]) in (2, 10)


@pytest.mark.parametrize("case", KATS["expectation"], ids=lambda c: c["name"])
def test_expectation_reference_results(gpu, case):
    labels, coeffs, psi = kd.campaign6_expectation(case)
    got = PauliSum.from_labels(labels, coeffs).to_device().expectation_statevector(psi)
    # Published values are MI300X device results; reference tolerance 1e-11.
    assert got.real == pytest.approx(case["result_real"], abs=1e-11)
    assert got.imag == pytest.approx(case["result_imag"], abs=1e-11)


def test_expectation_edges_and_errors(gpu):
    psi = random_statevector(3, 1)
    assert PauliSum.empty(3).to_device().expectation_statevector(psi) == 0
    ident = PauliSum.from_labels(["III"], [2.5])
    assert ident.to_device().expectation_statevector(psi) == pytest.approx(2.5, abs=1e-15)
    d = PauliSum.from_labels(["XYZ"]).to_device()
    with pytest.raises(ValueError, match="len\\(psi\\) == 2 \\*\\* num_qubits"):
        d.expectation_statevector(psi[:4])
    with pytest.raises(TypeError, match="complex64 or complex128"):
        d.expectation_statevector(psi.real.copy())
    with pytest.raises(ValueError, match="1-dimensional"):
        d.expectation_statevector(psi.reshape(2, 4))
    with pytest.raises(TypeError, match="C-contiguous"):
        d.expectation_statevector(np.concatenate([psi, psi])[::2])


@pytest.mark.parametrize("n", [1, 6, 11])
def test_adjoint_sigma_and_term_values(gpu, orc, n):
    import torch

    op = random_pauli_masks(n, 200, 0.4, 3 * n)
    psi = random_statevector(n, 4 * n)
    dop = op.to_device()
    t = torch.from_numpy(psi).to("cuda:0")
    sigma = torch.empty_like(t)
    sigma_dag = torch.empty_like(t)
    tv = torch.empty(op.num_terms, dtype=torch.complex128, device="cuda:0")
    e = dop.expectation_adjoint(t.data_ptr(), t.numel(), tv.data_ptr(), sigma.data_ptr(), sigma_dag.data_ptr())
    want_sigma = orc.apply(*packed(op), psi)
    cdag = packed(op)[:3] + (op.coeffs().conj(),)
    want_dag = orc.apply(*cdag, psi)
    s = scale(op, psi)
    np.testing.assert_allclose(sigma.cpu().numpy(), want_sigma, rtol=0, atol=RTOL * s)
    np.testing.assert_allclose(sigma_dag.cpu().numpy(), want_dag, rtol=0, atol=RTOL * s)
    want_tv = orc.term_values(n, op.x_words(), op.z_words(), psi)
    np.testing.assert_allclose(tv.cpu().numpy().real, want_tv, rtol=0, atol=RTOL)
    assert abs(e - orc.expectation_compensated(*packed(op), psi)) <= RTOL * s
    # <psi|sigma> = E
    assert abs(np.vdot(psi, sigma.cpu().numpy()) - e) <= RTOL * s


@pytest.mark.parametrize("gen,n,terms", [("random", 18, 300), ("chemistry", 19, 2000), ("diagonal", 18, 120)])
def test_tiled_adjoint_sigma_matches_oracle(gpu, orc, gen, n, terms):
    # States with n >= 18 take the tiled path: forward passes for E and the
    # per-term values, then sigma over the same passes (the chemistry case has
    # more than 32 diagonal keys, swept separately in the forward and carried
    # by the first pass in the sigma sweep).
    import torch

    from wolfgang_amd import PauliSum

    if gen == "chemistry":
        op = chemistry_masks(n, terms, 70 + n)
    elif gen == "random":
        op = random_pauli_masks(n, terms, 0.3, 71 + n)
    else:
        base = chemistry_masks(n, 4 * terms, 72 + n)
        keep = base.x_words()[:, 0] == 0
        op = PauliSum.from_packed(n, base.x_words()[keep][:terms], base.z_words()[keep][:terms],
                                  base.coeffs()[keep][:terms])
    psi = random_statevector(n, 73 + n)
    dop = op.to_device()
    assert dop.expectation_plan_info()["tiled"]
    t = torch.from_numpy(psi).to("cuda:0")
    rng = np.random.default_rng(n)
    coeffs = op.coeffs() * (1 + 0.5j * rng.normal(size=op.num_terms))  # complex override
    ct = torch.from_numpy(coeffs).to("cuda:0")
    sigma, sigma_dag = torch.empty_like(t), torch.empty_like(t)
    tv = torch.empty(op.num_terms, dtype=torch.complex128, device="cuda:0")
    for override in (None, ct):
        c = op.coeffs() if override is None else coeffs
        e = dop.expectation_adjoint(t.data_ptr(), t.numel(), tv.data_ptr(), sigma.data_ptr(),
                                    sigma_dag.data_ptr(),
                                    coeffs_ptr=None if override is None else override.data_ptr())
        want_sigma = orc.apply(n, op.x_words(), op.z_words(), c, psi, threads=8)
        want_dag = orc.apply(n, op.x_words(), op.z_words(), c.conj(), psi, threads=8)
        s = float(np.abs(c).sum()) * float(np.vdot(psi, psi).real)
        np.testing.assert_allclose(sigma.cpu().numpy(), want_sigma, rtol=0, atol=RTOL * s)
        np.testing.assert_allclose(sigma_dag.cpu().numpy(), want_dag, rtol=0, atol=RTOL * s)
        want_tv = orc.term_values(n, op.x_words(), op.z_words(), psi, threads=8)
        np.testing.assert_allclose(tv.cpu().numpy().real, want_tv, rtol=0, atol=RTOL)
        assert abs(e - np.sum(c * want_tv)) <= RTOL * s
        assert abs(np.vdot(psi, sigma.cpu().numpy()) - e) <= RTOL * s


@pytest.mark.parametrize("gen,n,terms", [("chemistry", 18, 2000), ("chemistry", 20, 3000),
                                         ("random", 18, 400), ("diagonal", 19, 300)])
def test_tiled_sweep_matches_oracle(gpu, orc, gen, n, terms):
    from wolfgang_amd import PauliSum

    if gen == "chemistry":
        op = chemistry_masks(n, terms, 40 + n)
    elif gen == "random":
        op = random_pauli_masks(n, terms, 0.3, 41 + n)
    else:
        r = random_pauli_masks(n, terms, 0.4, 42 + n)
        op = PauliSum.from_packed(n, np.zeros_like(r.x_words()), r.z_words(), r.coeffs())
    psi = random_statevector(n, 43 + n)
    dop = op.to_device()
    info = dop.expectation_plan_info()
    assert info["tiled"], info
    got = dop.expectation_statevector(psi)
    want = orc.expectation_compensated(*packed(op), psi, threads=8)
    assert abs(got - want) <= RTOL * scale(op, psi)
    assert dop.expectation_statevector(psi) == got


@pytest.mark.parametrize("variant", ["sigma_values", "pair_no_energy", "sigma_no_energy", "long_pass"])
def test_tiled_adjoint_variants(gpu, orc, variant):
    """The tiled adjoint in each of its forms: sigma with E and the per-term
    values (forward sweeps, then the sigma sweep), sigma and sigma^dagger with
    no energy wanted (the sigma sweep alone), sigma alone with no energy, and a
    pass of 300 keys on one x-mask. Repeated calls give identical bits."""
    import torch

    from test_sigma_plan import one_x_many_z

    n = 18
    op = one_x_many_z(n, 300, 84) if variant == "long_pass" else random_pauli_masks(n, 500, 0.3, 91)
    psi = random_statevector(n, 92)
    dop = op.to_device()
    assert dop.expectation_plan_info()["tiled"]
    t = torch.from_numpy(psi).to("cuda:0")
    sigma, sigma_dag = torch.empty_like(t), torch.empty_like(t)
    tv = torch.empty(op.num_terms, dtype=torch.complex128, device="cuda:0")
    pair = variant in ("pair_no_energy", "long_pass")
    values = variant in ("sigma_values", "long_pass")

    def call():
        return dop.expectation_adjoint(t.data_ptr(), t.numel(), tv.data_ptr() if values else None,
                                       sigma.data_ptr(), sigma_dag.data_ptr() if pair else None,
                                       want_energy=values)

    e = call()
    s = scale(op, psi)
    np.testing.assert_allclose(sigma.cpu().numpy(), orc.apply(*packed(op), psi, threads=8), rtol=0, atol=RTOL * s)
    if pair:
        want_dag = orc.apply(n, op.x_words(), op.z_words(), op.coeffs().conj(), psi, threads=8)
        np.testing.assert_allclose(sigma_dag.cpu().numpy(), want_dag, rtol=0, atol=RTOL * s)
    if values:
        want_tv = orc.term_values(n, op.x_words(), op.z_words(), psi, threads=8)
        np.testing.assert_allclose(tv.cpu().numpy().real, want_tv, rtol=0, atol=RTOL)
        assert abs(e - orc.expectation_compensated(*packed(op), psi, threads=8)) <= RTOL * s
    else:
        assert e is None
    first = (sigma.clone(), tv.clone())
    e2 = call()
    assert torch.equal(sigma, first[0]) and e2 == e
    if values:
        assert torch.equal(tv, first[1])
