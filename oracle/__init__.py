"""CPU oracle for the packed Pauli hot path: TEST INFRASTRUCTURE ONLY.

Python access to ``oracle/liboracle.so`` (the C++ restatement of the
reference's scalar path, wolfgang_oracle.cpp) and, where it was built, to
``oracle/_ref/libwolfgang_ref.so`` (the reference's own scalar sources compiled
in place by ``oracle/Makefile``). Only the test suite, the smoke entry point and
the benchmark's CPU-baseline leg import this package; the product library never
does.

Operators are passed as ``(num_qubits, x, z, coeffs)`` with x, z uint64 arrays
of shape (T, ceil(n/64)) and coeffs complex128 of shape (T,).
"""

from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ORACLE_LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libwolfgang_ref.so"

_vp = C.c_void_p
_sz = C.c_size_t


def _words(n: int) -> int:
    return 0 if n == 0 else (n + 63) // 64


def _arr(a: np.ndarray | None):
    if a is None or a.size == 0:
        return None
    return a.ctypes.data


def _prep(n: int, x, z, c):
    w = _words(n)
    c = np.ascontiguousarray(c, dtype=np.complex128).reshape(-1)
    t = c.shape[0]
    x = np.ascontiguousarray(x, dtype=np.uint64).reshape(t, w)
    z = np.ascontiguousarray(z, dtype=np.uint64).reshape(t, w)
    return x, z, c


def _z_count_args(bitstrings, counts):
    """Concatenated UTF-8 bitstrings with row offsets, and the counts."""
    enc = [b.encode("utf-8") for b in bitstrings]
    offsets = np.zeros(len(enc) + 1, dtype=np.uintp)
    if enc:
        offsets[1:] = np.cumsum([len(b) for b in enc])
    counts = np.ascontiguousarray(counts, dtype=np.float64)
    return b"".join(enc), offsets, counts


class _Lib:
    def __init__(self, path: pathlib.Path, prefix: str) -> None:
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle` / `make -C oracle ref`)")
        self.lib = C.CDLL(str(path))
        self.p = prefix
        getattr(self.lib, f"{prefix}_last_error").restype = C.c_char_p
        getattr(self.lib, f"{prefix}_result_terms").restype = _sz
        getattr(self.lib, f"{prefix}_result_terms").argtypes = [_vp]
        getattr(self.lib, f"{prefix}_result_copy").argtypes = [_vp, _vp, _vp, _vp]
        getattr(self.lib, f"{prefix}_result_free").argtypes = [_vp]

    def fn(self, name: str):
        return getattr(self.lib, f"{self.p}_{name}")

    def check(self, rc: int) -> None:
        if rc == 0:
            return
        msg = self.fn("last_error")().decode()
        if rc == 1:
            raise ValueError(msg)
        raise RuntimeError(msg)

    def result(self, handle: int, n: int):
        t = self.fn("result_terms")(handle)
        w = _words(n)
        x = np.zeros((t, w), dtype=np.uint64)
        z = np.zeros((t, w), dtype=np.uint64)
        c = np.zeros(t, dtype=np.complex128)
        self.fn("result_copy")(handle, _arr(x), _arr(z), _arr(c))
        self.fn("result_free")(handle)
        return n, x, z, c

    def simplify(self, n, x, z, c, atol=1e-12, rtol=0.0):
        x, z, c = _prep(n, x, z, c)
        out = _vp()
        f = self.fn("simplify")
        f.argtypes = [_sz, _sz, _vp, _vp, _vp, C.c_double, C.c_double, C.POINTER(_vp)]
        self.check(f(n, c.shape[0], _arr(x), _arr(z), _arr(c), atol, rtol, C.byref(out)))
        return self.result(out.value, n)

    def matmul(self, n, lhs, rhs, simplify=True, max_terms=50_000_000):
        lx, lz, lc = _prep(n, *lhs)
        rx, rz, rc = _prep(n, *rhs)
        out = _vp()
        f = self.fn("matmul")
        f.argtypes = [_sz, _sz, _vp, _vp, _vp, _sz, _vp, _vp, _vp, C.c_int, _sz, C.POINTER(_vp)]
        self.check(
            f(n, lc.shape[0], _arr(lx), _arr(lz), _arr(lc), rc.shape[0], _arr(rx), _arr(rz), _arr(rc),
              int(simplify), max_terms, C.byref(out))
        )
        return self.result(out.value, n)

    def expectation(self, n, x, z, c, psi):
        x, z, c = _prep(n, x, z, c)
        out = (C.c_double * 2)()
        if psi.dtype == np.complex64:
            psi = np.ascontiguousarray(psi)
            f = self.fn("expectation_c64")
        else:
            psi = np.ascontiguousarray(psi, dtype=np.complex128)
            f = self.fn("expectation")
        f.argtypes = [_sz, _sz, _vp, _vp, _vp, _vp, _sz, C.POINTER(C.c_double)]
        self.check(f(n, c.shape[0], _arr(x), _arr(z), _arr(c), _arr(psi), psi.shape[0], out))
        return complex(out[0], out[1])


class Oracle(_Lib):
    """The C++ restatement (wolfgang_oracle.cpp)."""

    def __init__(self) -> None:
        super().__init__(ORACLE_LIB, "orc")

    def commutes_with(self, n, lhs_xz, rhs_xz, max_entries=100_000_000, variant=0, threads=1):
        """variant: 0 scalar, 1 AVX-512 VPOPCNTDQ, 2 row-partitioned threads."""
        lx, lz = (np.ascontiguousarray(a, dtype=np.uint64) for a in lhs_xz)
        rx, rz = (np.ascontiguousarray(a, dtype=np.uint64) for a in rhs_xz)
        L, R = lx.shape[0], rx.shape[0]
        out = np.zeros(L * R, dtype=np.uint8)
        f = self.fn("commutes_with")
        f.argtypes = [_sz, _sz, _vp, _vp, _sz, _vp, _vp, _sz, C.c_int, C.c_int, _vp]
        self.check(f(n, L, _arr(lx), _arr(lz), R, _arr(rx), _arr(rz), max_entries, variant, threads, _arr(out)))
        return out.reshape(L, R)

    def set_threads(self, threads: int) -> None:
        """Worker threads for simplify's sort and matmul's products (the
        results are the same for any count)."""
        self.fn("set_threads").argtypes = [C.c_int]
        self.fn("set_threads")(int(threads))

    def has_avx512(self) -> bool:
        return bool(self.fn("has_avx512_popcnt")())

    def group_commuting(self, n, x, z, mode):
        """Greedy largest-first grouping; returns a list of member-index arrays
        (groups in order, members in placement order)."""
        x = np.ascontiguousarray(x, dtype=np.uint64)
        z = np.ascontiguousarray(z, dtype=np.uint64)
        T = x.shape[0]
        sizes = np.zeros(max(T, 1), dtype=np.uint32)
        members = np.zeros(max(T, 1), dtype=np.uint32)
        g = C.c_size_t()
        f = self.fn("group_commuting")
        f.argtypes = [_sz, _sz, _vp, _vp, C.c_int, C.POINTER(C.c_size_t), _vp, _vp]
        self.check(f(n, T, _arr(x), _arr(z), {"qwc": 0, "full": 1}[mode], C.byref(g), _arr(sizes), _arr(members)))
        out, k = [], 0
        for i in range(g.value):
            out.append(members[k:k + sizes[i]].astype(np.int64))
            k += int(sizes[i])
        return out

    def simplify_hash_path(self, n, x, z) -> bool:
        x = np.ascontiguousarray(x, dtype=np.uint64)
        z = np.ascontiguousarray(z, dtype=np.uint64)
        f = self.fn("simplify_hash_path")
        f.argtypes = [_sz, _sz, _vp, _vp]
        return bool(f(n, x.shape[0], _arr(x), _arr(z)))

    def term_values(self, n, x, z, psi, threads=1):
        x = np.ascontiguousarray(x, dtype=np.uint64)
        z = np.ascontiguousarray(z, dtype=np.uint64)
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        out = np.zeros(x.shape[0], dtype=np.float64)
        f = self.fn("term_values")
        f.argtypes = [_sz, _sz, _vp, _vp, _vp, _sz, C.c_int, _vp]
        self.check(f(n, x.shape[0], _arr(x), _arr(z), _arr(psi), psi.shape[0], threads, _arr(out)))
        return out

    def expectation_compensated(self, n, x, z, c, psi, threads=1):
        x, z, c = _prep(n, x, z, c)
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        out = (C.c_double * 2)()
        f = self.fn("expectation_compensated")
        f.argtypes = [_sz, _sz, _vp, _vp, _vp, _vp, _sz, C.c_int, C.POINTER(C.c_double)]
        self.check(f(n, c.shape[0], _arr(x), _arr(z), _arr(c), _arr(psi), psi.shape[0], threads, out))
        return complex(out[0], out[1])

    def expectation_z_counts(self, n, x, z, c, bitstrings, counts):
        """E over computational-basis counts (diagonal terms only)."""
        x, z, c = _prep(n, x, z, c)
        chars, offsets, counts = _z_count_args(bitstrings, counts)
        out = (C.c_double * 2)()
        f = self.fn("expectation_z_counts")
        f.argtypes = [_sz, _sz, _vp, _vp, _vp, C.c_char_p, _vp, _vp, _sz, C.POINTER(C.c_double)]
        self.check(f(n, c.shape[0], _arr(x), _arr(z), _arr(c), chars, _arr(offsets), _arr(counts),
                     counts.shape[0], out))
        return complex(out[0], out[1])

    def apply(self, n, x, z, c, psi, threads=1):
        """sigma = H psi."""
        x, z, c = _prep(n, x, z, c)
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        out = np.zeros_like(psi)
        f = self.fn("apply")
        f.argtypes = [_sz, _sz, _vp, _vp, _vp, _vp, _sz, C.c_int, _vp]
        self.check(f(n, c.shape[0], _arr(x), _arr(z), _arr(c), _arr(psi), psi.shape[0], threads, _arr(out)))
        return out


    def commutes_packed(self, n, lhs_xz, rhs_xz, threads=1):
        """Packed commutation bits (rows x ceil(R/64) uint64 words) from the
        row-threaded kernel."""
        lx, lz = (np.ascontiguousarray(a, dtype=np.uint64) for a in lhs_xz)
        rx, rz = (np.ascontiguousarray(a, dtype=np.uint64) for a in rhs_xz)
        L, R = lx.shape[0], rx.shape[0]
        out = np.zeros((L, (R + 63) // 64), dtype=np.uint64)
        f = self.fn("commutes_packed")
        f.argtypes = [_sz, _sz, _vp, _vp, _sz, _vp, _vp, C.c_int, _vp]
        self.check(f(n, L, _arr(lx), _arr(lz), R, _arr(rx), _arr(rz), threads, _arr(out)))
        return out

    def apply_rows(self, n, x, z, c, psi, rows, threads=1):
        """sigma = H psi at the given output rows."""
        x, z, c = _prep(n, x, z, c)
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        out = np.zeros(rows.shape[0], dtype=np.complex128)
        f = self.fn("apply_rows")
        f.argtypes = [_sz, _sz, _vp, _vp, _vp, _vp, _sz, _vp, _sz, C.c_int, _vp]
        self.check(f(n, c.shape[0], _arr(x), _arr(z), _arr(c), _arr(psi), psi.shape[0], _arr(rows),
                     rows.shape[0], threads, _arr(out)))
        return out

    @staticmethod
    def _word0(n, x, z):
        """Word 0 of one-word masks as flat arrays (zeros at n = 0)."""
        x = np.asarray(x, dtype=np.uint64)
        z = np.asarray(z, dtype=np.uint64)
        T = x.shape[0]
        if n == 0:
            return np.zeros(T, dtype=np.uint64), np.zeros(T, dtype=np.uint64)
        return (np.ascontiguousarray(x.reshape(T, -1)[:, 0]), np.ascontiguousarray(z.reshape(T, -1)[:, 0]))

    def distinct_keys(self, n, x, z):
        """(number of distinct keys, key id of each term) in first-appearance order."""
        x, z = self._word0(n, x, z)
        T = x.shape[0]
        key = np.zeros(max(T, 1), dtype=np.uint32)
        k = C.c_size_t()
        f = self.fn("distinct_keys")
        f.argtypes = [_sz, _sz, _vp, _vp, _vp, C.POINTER(C.c_size_t)]
        self.check(f(n, T, _arr(x), _arr(z), _arr(key), C.byref(k)))
        return k.value, key[:T]

    def key_values_grouped(self, n, x, z, psi, threads=1, progress=None, pieces=1):
        """Per-key <P_k> (keys in first-appearance order) by the x-grouped,
        chunk-paired double-double evaluation; `progress(done, total)` is
        called after each of `pieces` chunk ranges."""
        x, z = self._word0(n, x, z)
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        K, _ = self.distinct_keys(n, x, z)
        hi = np.zeros(max(K, 1), dtype=np.float64)
        lo = np.zeros(max(K, 1), dtype=np.float64)
        chunks = psi.shape[0] >> min(n, 16)
        f = self.fn("key_values_accumulate")
        f.argtypes = [_sz, _sz, _vp, _vp, _vp, _sz, _sz, _sz, C.c_int, _vp, _vp]
        bounds = np.linspace(0, chunks, max(1, min(pieces, chunks)) + 1).astype(np.int64)
        for i in range(len(bounds) - 1):
            self.check(f(n, x.shape[0], _arr(x), _arr(z), _arr(psi), psi.shape[0], int(bounds[i]),
                         int(bounds[i + 1]), threads, _arr(hi), _arr(lo)))
            if progress is not None:
                progress(int(bounds[i + 1]), chunks)
        return (hi + lo)[:K]

    def expectation_grouped(self, n, x, z, c, psi, threads=1, progress=None, pieces=1):
        """E = sum_t c_t <P_t> from key_values_grouped, the products summed
        exactly rounded (math.fsum)."""
        import math

        x, z, c = _prep(n, x, z, c)
        _, key = self.distinct_keys(n, x, z)
        vals = self.key_values_grouped(n, x, z, psi, threads, progress, pieces)
        tv = vals[key] if c.shape[0] else np.zeros(0)
        return complex(math.fsum((c.real * tv).tolist()), math.fsum((c.imag * tv).tolist()))


class Reference(_Lib):
    """The reference's own scalar sources, compiled in place (oracle/_ref)."""

    def __init__(self) -> None:
        super().__init__(REF_LIB, "wgref")

    def commutes_with(self, n, lhs_xz, rhs_xz, max_entries=100_000_000):
        lx, lz = (np.ascontiguousarray(a, dtype=np.uint64) for a in lhs_xz)
        rx, rz = (np.ascontiguousarray(a, dtype=np.uint64) for a in rhs_xz)
        L, R = lx.shape[0], rx.shape[0]
        out = np.zeros(L * R, dtype=np.uint8)
        f = self.fn("commutes_with")
        f.argtypes = [_sz, _sz, _vp, _vp, _sz, _vp, _vp, _sz, _vp]
        self.check(f(n, L, _arr(lx), _arr(lz), R, _arr(rx), _arr(rz), max_entries, _arr(out)))
        return out.reshape(L, R)

    def group_commuting(self, n, x, z, c, mode, max_terms_for_graph=50_000):
        """The reference's groups as a list of (x, z, c) arrays."""
        x, z, c = _prep(n, x, z, c)
        T = c.shape[0]
        W = x.shape[1] if x.ndim == 2 else 0
        sizes = np.zeros(max(T, 1), dtype=np.uint64)
        gx = np.zeros((max(T, 1), max(W, 1)), dtype=np.uint64)
        gz = np.zeros_like(gx)
        gc = np.zeros(max(T, 1), dtype=np.complex128)
        g = C.c_size_t()
        f = self.fn("group_commuting")
        f.argtypes = [_sz, _sz, _vp, _vp, _vp, C.c_char_p, _sz, C.POINTER(C.c_size_t), _vp, _vp, _vp, _vp]
        self.check(f(n, T, _arr(x), _arr(z), _arr(c), mode.encode(), max_terms_for_graph, C.byref(g),
                     _arr(sizes), _arr(gx), _arr(gz), _arr(gc)))
        out, k = [], 0
        for i in range(g.value):
            m = int(sizes[i])
            out.append((gx[k:k + m, :W], gz[k:k + m, :W], gc[k:k + m]))
            k += m
        return out

    def expectation_z_counts(self, n, x, z, c, bitstrings, counts):
        """E over computational-basis counts (diagonal terms only)."""
        x, z, c = _prep(n, x, z, c)
        chars, offsets, counts = _z_count_args(bitstrings, counts)
        out = (C.c_double * 2)()
        f = self.fn("expectation_z_counts")
        f.argtypes = [_sz, _sz, _vp, _vp, _vp, C.c_char_p, _vp, _vp, _sz, C.POINTER(C.c_double)]
        self.check(f(n, c.shape[0], _arr(x), _arr(z), _arr(c), chars, _arr(offsets), _arr(counts),
                     counts.shape[0], out))
        return complex(out[0], out[1])


_oracle: Oracle | None = None
_reference: Reference | None = None


def oracle() -> Oracle:
    global _oracle
    if _oracle is None:
        _oracle = Oracle()
    return _oracle


def reference() -> Reference | None:
    """The compiled reference, or None where oracle/_ref was not built."""
    global _reference
    if _reference is None and REF_LIB.exists():
        _reference = Reference()
    return _reference
