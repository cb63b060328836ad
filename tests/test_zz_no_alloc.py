"""The allocation-free expectation entry point (wg_expectation_noalloc),
called through the C ABI: its energies must equal wg_expectation's bit for bit
(same plan, same kernels, same fixed-order reductions) on two calls in a row,
the second with a smaller operator and state, so the shared static scratch
holds the first call's contents when the second runs. Also its error paths:
no plan built yet, and a plan whose partial sums exceed the scratch."""

import ctypes as C

import numpy as np
import pytest

from wolfgang_amd import _native
from wolfgang_amd.workloads import chemistry_masks, random_pauli_masks, random_statevector_device

pytestmark = pytest.mark.gpu


def _noalloc(dop, psi, out):
    lib = _native.load()
    stream = C.c_void_p(None)
    return lib.wg_expectation_noalloc(C.c_void_p(dop._handle()), C.c_void_p(psi.data_ptr()),
                                      psi.numel(), C.c_void_p(out.data_ptr()), stream)


def test_noalloc_matches_wg_expectation_two_calls(gpu):
    import torch

    cases = [
        (chemistry_masks(20, 3000, 71), random_statevector_device(20, 72, 0)),   # tiled plan
        (random_pauli_masks(13, 900, 0.4, 73), random_statevector_device(13, 74, 0)),  # direct plan
    ]
    out = torch.zeros(2, dtype=torch.complex128, device="cuda:0")
    for op, psi in cases:
        dop = op.to_device()
        want = dop.expectation_statevector(psi)  # builds and caches the plan
        print(f"[no-alloc] n={op.num_qubits} T={op.num_terms} plan={dop.expectation_plan_info()}", flush=True)
        out.zero_()
        torch.cuda.synchronize()
        assert _noalloc(dop, psi, out[1:]) == 0, _native.last_error()
        torch.cuda.synchronize()
        got = complex(out[1].item())
        assert got == want, (got, want)
        assert complex(out[0].item()) == 0  # nothing written outside the output


def test_noalloc_requires_a_built_plan(gpu):
    import torch

    op = random_pauli_masks(10, 50, 0.4, 75).to_device()
    psi = random_statevector_device(10, 76, 0)
    out = torch.zeros(1, dtype=torch.complex128, device="cuda:0")
    assert _noalloc(op, psi, out) == 1
    assert "plan" in _native.last_error()
    op.expectation_plan_info()
    assert _noalloc(op, psi, out) == 0, _native.last_error()
    torch.cuda.synchronize()
    assert complex(out[0].item()) == op.expectation_statevector(psi)


def test_noalloc_rejects_plans_beyond_the_scratch(gpu):
    import torch

    # 24 qubits, 40,000 random terms: ~26k x-groups x 64 tile blocks of partial
    # rows (128 B each) -- more than the 256 MiB scratch.
    op = random_pauli_masks(24, 40_000, 0.25, 77).to_device()
    op.expectation_plan_info()
    psi = random_statevector_device(24, 78, 0)
    out = torch.zeros(1, dtype=torch.complex128, device="cuda:0")
    assert _noalloc(op, psi, out) == 1
    assert "maximum" in _native.last_error()
