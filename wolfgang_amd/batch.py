"""Batched forward + adjoint over many (psi, H) pairs in one graph replay.

``BatchedExpectation(ops, psis)`` builds the batch (see include/wolfgang_amd.h,
wg_batch_create): small instances as units of one persistent, queue-driven
kernel, large (tiled) instances as tiled sweeps on worker streams, both captured
into one HIP graph; ``run()`` replays it on the current torch stream. Outputs
per instance: E_i (``energies``), per-term values <P_k> and sigma_i = H_i psi_i
(and H^dagger psi when asked). The coefficient tensors are read on every run.

The tunables (forward chunk, sigma chunk, persistent grid, tiled worker
streams, tiled part inside the graph) are found by timing the actual batch once
and remembered in a JSON cache keyed by the device and the batch's shape
signature (``$WOLFGANG_AMD_CACHE`` or ``~/.cache/wolfgang_amd/autotune.json``).

``plain_forward_adjoint`` is the per-instance path over the same outputs (one
expectation_adjoint call per instance), the baseline the batched launch is
measured against.
"""

from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import pathlib
from typing import Sequence

import torch

from . import _native
from .pauli_sum import DevicePauliSum


def _cache_path() -> pathlib.Path:
    env = os.environ.get("WOLFGANG_AMD_CACHE")
    if env:
        return pathlib.Path(env)
    return pathlib.Path.home() / ".cache" / "wolfgang_amd" / "autotune.json"


def _load_cache() -> dict:
    try:
        return json.loads(_cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store_cache(key: str, value: list[int]) -> None:
    cache = _load_cache()
    cache[key] = value
    path = _cache_path()
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(cache, indent=1, sort_keys=True))
        tmp.replace(path)
    except OSError:
        pass


def _ptr_array(values: Sequence[int | None]):
    arr = (C.c_void_p * len(values))()
    for i, v in enumerate(values):
        arr[i] = v
    return arr


class BatchedExpectation:
    def __init__(
        self,
        ops: Sequence[DevicePauliSum],
        psis: Sequence[torch.Tensor],
        coeffs: Sequence[torch.Tensor | None] | None = None,
        term_values: bool = True,
        sigma: bool = True,
        sigma_dagger: bool = False,
        tunables: Sequence[int] | None = None,
        use_cache: bool = True,
    ) -> None:
        if len(ops) != len(psis) or not ops:
            raise ValueError("ops and psis must be non-empty and of equal length")
        device = ops[0].device
        self.ops = list(ops)
        self.psis = []
        for op, psi in zip(ops, psis):
            if not (isinstance(psi, torch.Tensor) and psi.is_cuda and psi.dtype == torch.complex128):
                raise TypeError("batched statevectors must be complex128 device tensors")
            if psi.numel() != 1 << op.num_qubits:
                raise ValueError("expectation_statevector requires len(psi) == 2 ** num_qubits")
            self.psis.append(psi.contiguous())
        dev = torch.device("cuda", device)
        self.coeffs = [None] * len(ops)
        if coeffs is not None:
            self.coeffs = [c.to(dev, torch.complex128).contiguous() if c is not None else None for c in coeffs]
        self.term_values = [torch.empty(op.num_terms, dtype=torch.complex128, device=dev) for op in ops] if term_values else None
        self.sigma = [torch.empty_like(p) for p in self.psis] if sigma else None
        self.sigma_dagger = [torch.empty_like(p) for p in self.psis] if sigma_dagger else None
        self.energies = torch.zeros(len(ops), dtype=torch.complex128, device=dev)
        sig = hashlib.sha256(
            json.dumps([[op.num_qubits, op.num_terms] for op in ops]).encode()
            + torch.cuda.get_device_name(device).encode()
            + bytes([int(sigma), int(sigma_dagger), int(term_values)])
        ).hexdigest()[:24]
        self.cache_key = f"batch:{sig}"
        if tunables is None and use_cache:
            cached = _load_cache().get(self.cache_key)
            if cached and len(cached) == 5:
                tunables = cached
        if tunables is not None and len(tunables) != 5:
            raise ValueError("batch tunables are 5 integers")
        tun = (C.c_int * 5)(*tunables) if tunables is not None else None
        handle = C.c_void_p()
        stream = torch.cuda.current_stream(dev).cuda_stream
        _native.call(
            "wg_batch_create",
            len(ops),
            _ptr_array([op._handle() for op in ops]),
            _ptr_array([p.data_ptr() for p in self.psis]),
            _ptr_array([c.data_ptr() if c is not None else None for c in self.coeffs]),
            _ptr_array([t.data_ptr() for t in self.term_values]) if self.term_values else None,
            _ptr_array([s.data_ptr() for s in self.sigma]) if self.sigma else None,
            _ptr_array([s.data_ptr() for s in self.sigma_dagger]) if self.sigma_dagger else None,
            self.energies.data_ptr(),
            tun,
            stream,
            C.byref(handle),
        )
        self._h = handle.value
        self._device = dev
        info = (C.c_int * 5)()
        units, items, tiled = C.c_size_t(), C.c_size_t(), C.c_size_t()
        _native.call("wg_batch_info", self._h, info, C.byref(units), C.byref(items))
        _native.call("wg_batch_tiled_count", self._h, C.byref(tiled))
        self.tunables = [int(v) for v in info]
        self.units = units.value
        self.items = items.value
        self.tiled_instances = tiled.value
        if use_cache and tunables is None:
            _store_cache(self.cache_key, self.tunables)

    def run(self) -> None:
        """Enqueue one replay of the captured graph on the current stream."""
        _native.call("wg_batch_run", self._h, torch.cuda.current_stream(self._device).cuda_stream)

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            try:
                _native.load().wg_batch_destroy(h)
            except Exception:
                pass
            self._h = None


def plain_forward_adjoint(batch: BatchedExpectation) -> None:
    """The same outputs computed instance by instance (one synchronous
    expectation_adjoint call each) on the batch's buffers."""
    stream = torch.cuda.current_stream(batch._device).cuda_stream
    energies = []
    for i, (op, psi) in enumerate(zip(batch.ops, batch.psis)):
        e = op.expectation_adjoint(
            psi.data_ptr(), psi.numel(),
            batch.term_values[i].data_ptr() if batch.term_values else None,
            batch.sigma[i].data_ptr() if batch.sigma else None,
            batch.sigma_dagger[i].data_ptr() if batch.sigma_dagger else None,
            stream=stream,
            coeffs_ptr=batch.coeffs[i].data_ptr() if batch.coeffs[i] is not None else None,
        )
        energies.append(e)
    batch.energies.copy_(torch.tensor(energies, dtype=torch.complex128))


class ProductBatch:
    """Many Hamiltonian products ``lhs[i] @ rhs[i]`` through one persistent,
    queue-fed product kernel per chunk (captured in a HIP graph), then each
    pair's simplify. ``run()`` returns one new ``DevicePauliSum`` per pair,
    bit-identical to ``lhs[i].matmul(rhs[i], simplify)``. The operands must
    stay alive while the batch exists."""

    def __init__(
        self,
        lhs: Sequence[DevicePauliSum],
        rhs: Sequence[DevicePauliSum],
        max_intermediate_terms: int = 50_000_000,
        budget_bytes: int = 0,
    ) -> None:
        if len(lhs) != len(rhs) or not lhs:
            raise ValueError("lhs and rhs must be non-empty and of equal length")
        self.lhs, self.rhs = list(lhs), list(rhs)
        handle = C.c_void_p()
        stream = torch.cuda.current_stream(torch.device("cuda", lhs[0].device)).cuda_stream
        _native.call(
            "wg_product_batch_create", len(lhs), _ptr_array([a._handle() for a in self.lhs]),
            _ptr_array([b._handle() for b in self.rhs]), int(max_intermediate_terms),
            int(budget_bytes), stream, C.byref(handle),
        )
        self._h = handle.value
        self._device = lhs[0].device
        pairs, chunks, total = C.c_size_t(), C.c_size_t(), C.c_size_t()
        _native.call("wg_product_batch_info", self._h, C.byref(pairs), C.byref(chunks), C.byref(total))
        self.chunks = chunks.value
        self.total_products = total.value

    def run(self, simplify: bool = True) -> list[DevicePauliSum]:
        outs = (C.c_void_p * len(self.lhs))()
        stream = torch.cuda.current_stream(torch.device("cuda", self._device)).cuda_stream
        _native.call("wg_product_batch_run", self._h, int(bool(simplify)), stream, outs)
        return [DevicePauliSum._wrap(outs[i]) for i in range(len(self.lhs))]

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            try:
                _native.load().wg_product_batch_destroy(h)
            except Exception:
                pass
            self._h = None


def plain_products(lhs: Sequence[DevicePauliSum], rhs: Sequence[DevicePauliSum],
                   simplify: bool = True, max_intermediate_terms: int = 50_000_000) -> list[DevicePauliSum]:
    """The same products pair by pair (one synchronous matmul call each)."""
    return [a.matmul(b, simplify=simplify, max_intermediate_terms=max_intermediate_terms)
            for a, b in zip(lhs, rhs)]
