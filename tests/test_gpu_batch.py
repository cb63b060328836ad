"""Batched forward + adjoint (one graph replay: queued small instances in a
persistent kernel, tiled instances on worker streams) against the
per-instance path and the oracle (1e-12 relative to sum|c| ||psi||^2)."""

import numpy as np
import pytest

from conftest import packed
from wolfgang_amd.workloads import batch_shapes, chemistry_masks, random_pauli_masks, random_statevector

pytestmark = pytest.mark.gpu


def make_batch(count, seed, n_range=(3, 11), t_range=(5, 300)):
    shapes = batch_shapes(count, seed, n_range, t_range)
    ops, psis = [], []
    for i, (n, t) in enumerate(shapes):
        op = random_pauli_masks(n, t, 0.25, seed + i) if i % 2 else chemistry_masks(max(n, 2), t, seed + i)
        ops.append(op)
        psis.append(random_statevector(op.num_qubits, seed + 1000 + i))
    return ops, psis


@pytest.mark.parametrize("tunables", [None, (10, 256, 64, 1, 1)])
def test_batch_matches_oracle_and_plain_path(gpu, orc, tunables, tmp_path, monkeypatch):
    import torch

    from wolfgang_amd.batch import BatchedExpectation, plain_forward_adjoint

    monkeypatch.setenv("WOLFGANG_AMD_CACHE", str(tmp_path / "tune.json"))
    ops, psis = make_batch(24, 77)
    dops = [op.to_device() for op in ops]
    tpsi = [torch.from_numpy(p).to("cuda:0") for p in psis]
    b = BatchedExpectation(dops, tpsi, sigma=True, sigma_dagger=True, tunables=tunables)
    assert b.units > 0
    for _ in range(3):  # replays with eager work in between
        b.run()
        torch.cuda.synchronize()
        energies = b.energies.cpu().numpy()
        for i, (op, psi) in enumerate(zip(ops, psis)):
            scale = float(np.abs(op.coeffs()).sum())
            want = orc.expectation_compensated(*packed(op), psi)
            assert abs(energies[i] - want) <= 1e-12 * scale
            np.testing.assert_allclose(b.sigma[i].cpu().numpy(), orc.apply(*packed(op), psi),
                                       rtol=0, atol=1e-12 * scale)
            tv = orc.term_values(op.num_qubits, op.x_words(), op.z_words(), psi)
            np.testing.assert_allclose(b.term_values[i].cpu().numpy().real, tv, rtol=0, atol=1e-12)
        _ = torch.randn(100, device="cuda:0").sum().item()
    sig_batch = [s.clone() for s in b.sigma]
    e_batch = b.energies.clone()
    plain_forward_adjoint(b)
    torch.cuda.synchronize()
    for i, op in enumerate(ops):
        scale = float(np.abs(op.coeffs()).sum())
        assert abs(complex(b.energies[i].item()) - complex(e_batch[i].item())) <= 1e-12 * scale
        assert torch.allclose(b.sigma[i], sig_batch[i], rtol=0, atol=1e-12 * scale)


def test_batch_with_coefficient_overrides(gpu, orc):
    import torch

    from wolfgang_amd.batch import BatchedExpectation

    ops, psis = make_batch(6, 91)
    rng = np.random.default_rng(4)
    new_c = [rng.normal(size=op.num_terms) + 1j * rng.normal(size=op.num_terms) for op in ops]
    b = BatchedExpectation([op.to_device() for op in ops], [torch.from_numpy(p).to("cuda:0") for p in psis],
                           coeffs=[torch.from_numpy(c) for c in new_c], sigma=True, use_cache=False)
    b.run()
    torch.cuda.synchronize()
    for i, (op, psi) in enumerate(zip(ops, psis)):
        scale = float(np.abs(new_c[i]).sum())
        want = orc.expectation_compensated(op.num_qubits, op.x_words(), op.z_words(), new_c[i], psi)
        assert abs(complex(b.energies[i].item()) - want) <= 1e-12 * scale
        np.testing.assert_allclose(
            b.sigma[i].cpu().numpy(), orc.apply(op.num_qubits, op.x_words(), op.z_words(), new_c[i], psi),
            rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("all_tiled,tunables", [(False, None), (True, None), (False, (12, 1024, 256, 2, 1)),
                                                 (False, (12, 1024, 256, 4, 0))])
def test_batch_with_tiled_instances(gpu, orc, all_tiled, tunables):
    # Instances with a tiled plan (18 qubits and up) run the tiled sweeps on
    # worker streams (inside the graph, or enqueued per run); the rest stay
    # queued. The coefficient tensors are read on every run: changing them in
    # place between runs changes E, sigma and sigma^dagger of every instance.
    import torch

    from wolfgang_amd.batch import BatchedExpectation, plain_forward_adjoint

    shapes = [(18, 400), (19, 150), (18, 90)] + ([] if all_tiled else [(9, 60), (12, 500), (5, 20)])
    ops = [random_pauli_masks(n, t, 0.25, 300 + i) if i % 2 else chemistry_masks(n, t, 300 + i)
           for i, (n, t) in enumerate(shapes)]
    psis = [random_statevector(n, 400 + i) for i, (n, _) in enumerate(shapes)]
    rng = np.random.default_rng(9)
    new_c = [rng.normal(size=op.num_terms) + 1j * rng.normal(size=op.num_terms) for op in ops]
    b = BatchedExpectation([op.to_device() for op in ops], [torch.from_numpy(p).to("cuda:0") for p in psis],
                           coeffs=[torch.from_numpy(c) for c in new_c], sigma=True, sigma_dagger=True,
                           tunables=tunables, use_cache=False)
    assert (b.units == 0) == all_tiled
    assert b.tiled_instances == 3
    if tunables is not None:
        assert b.tunables == list(tunables)
    for rnd in range(3):
        if rnd == 2:
            # New coefficients written into the same tensors.
            new_c = [rng.normal(size=op.num_terms) + 1j * rng.normal(size=op.num_terms) for op in ops]
            for t, c in zip(b.coeffs, new_c):
                t.copy_(torch.from_numpy(c))
        b.run()
        torch.cuda.synchronize()
        for i, (op, psi) in enumerate(zip(ops, psis)):
            n, x, z = op.num_qubits, op.x_words(), op.z_words()
            scale = float(np.abs(new_c[i]).sum())
            want = orc.expectation_compensated(n, x, z, new_c[i], psi)
            assert abs(complex(b.energies[i].item()) - want) <= 1e-12 * scale, (rnd, i)
            np.testing.assert_allclose(b.sigma[i].cpu().numpy(), orc.apply(n, x, z, new_c[i], psi, threads=8),
                                       rtol=0, atol=1e-12 * scale)
            np.testing.assert_allclose(b.sigma_dagger[i].cpu().numpy(),
                                       orc.apply(n, x, z, new_c[i].conj(), psi, threads=8),
                                       rtol=0, atol=1e-12 * scale)
            tv = orc.term_values(n, x, z, psi)
            np.testing.assert_allclose(b.term_values[i].cpu().numpy().real, tv, rtol=0, atol=1e-12)
        _ = torch.randn(100, device="cuda:0").sum().item()
    e_batch = b.energies.clone()
    plain_forward_adjoint(b)
    torch.cuda.synchronize()
    for i in range(len(ops)):
        scale = float(np.abs(new_c[i]).sum())
        assert abs(complex(b.energies[i].item()) - complex(e_batch[i].item())) <= 1e-12 * scale


# ------------------------------------------------------------ batched products

def make_product_pairs(seed, count=9):
    from wolfgang_amd.workloads import product_shapes

    pairs = []
    for i, (n, tl, tr) in enumerate(product_shapes(count, seed, t_range=(1, 400), n_range=(1, 256))):
        a = random_pauli_masks(n, tl, 0.25, seed + 3 * i)
        b = random_pauli_masks(n, tr, 0.25, seed + 3 * i + 1)
        pairs.append((a, b))
    # Duplicate-heavy operand (matmul's hash seed) and a zero-term pair.
    pool = random_pauli_masks(70, 5, 0.3, seed)
    idx = np.arange(4096) % 5
    dup = type(pool).from_packed(70, pool.x_words()[idx], pool.z_words()[idx],
                                 np.arange(4096, dtype=float) % 7 - 3)
    pairs.append((dup, random_pauli_masks(70, 3, 0.3, seed + 99)))
    pairs.append((random_pauli_masks(40, 0, 0.3, seed), random_pauli_masks(40, 7, 0.3, seed + 1)))
    return pairs


@pytest.mark.parametrize("budget", [0, 200_000])
def test_product_batch_bit_exact_vs_matmul(gpu, orc, budget):
    from conftest import assert_bit_exact
    from wolfgang_amd.batch import ProductBatch

    pairs = make_product_pairs(4242)
    lhs = [a.to_device() for a, _ in pairs]
    rhs = [b.to_device() for _, b in pairs]
    pb = ProductBatch(lhs, rhs, budget_bytes=budget)
    if budget:
        assert pb.chunks > 1
    for simplify in (True, False, True):  # repeated graph replays
        outs = pb.run(simplify=simplify)
        for (a, b), dl, dr, got in zip(pairs, lhs, rhs, outs):
            host = got.to_host()
            assert_bit_exact(host, packed(dl.matmul(dr, simplify=simplify).to_host()))
            assert_bit_exact(host, orc.matmul(a.num_qubits, packed(a)[1:], packed(b)[1:], simplify))


def test_product_batch_errors(gpu):
    from wolfgang_amd.batch import ProductBatch

    a = random_pauli_masks(8, 10, 0.3, 1).to_device()
    b = random_pauli_masks(9, 10, 0.3, 2).to_device()
    with pytest.raises(ValueError, match="same num_qubits"):
        ProductBatch([a], [b])
    c = random_pauli_masks(8, 10, 0.3, 3).to_device()
    with pytest.raises(ValueError, match="exceeds max_intermediate_terms"):
        ProductBatch([a], [c], max_intermediate_terms=99)
    with pytest.raises(ValueError):
        ProductBatch([], [])
