"""Build recipe for the native pieces, run in-tree.

* ``wolfgang_amd/libwolfgang_amd.so``: every ``wolfgang_amd/csrc/*.hip`` compiled
  for gfx950 with hipcc and linked into one shared library exporting the C ABI
  declared in ``include/wolfgang_amd.h``.
* ``oracle/liboracle.so`` (the CPU oracle; test infrastructure) and, where the
  reference tree is mounted, ``oracle/_ref/libwolfgang_ref.so``: both through
  ``oracle/Makefile``.

Usage: ``python -m wolfgang_amd.build_native`` (or ``__graft_entry__.build()``).
"""

from __future__ import annotations

import concurrent.futures
import os
import pathlib
import shutil
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parent.parent
PKG = ROOT / "wolfgang_amd"
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libwolfgang_amd.so"

ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    # The multiply kernel must round (ac - bd, ad + bc) like the x86-64
    # reference; kernels that want fused multiply-adds call fma() explicitly.
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and pathlib.Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the wolfgang_amd HIP library cannot be built")


def _compile(src: pathlib.Path) -> pathlib.Path:
    out = OBJ / (src.stem + ".o")
    cmd = [hipcc(), *HIP_FLAGS, "-c", str(src), "-o", str(out)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{proc.stderr}")
    return out


def build_library(jobs: int | None = None) -> pathlib.Path:
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(len(sources), max(1, (os.cpu_count() or 4)), 16)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objects)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stderr}")
    tmp.replace(LIB)
    return LIB


def build_oracle(reference: str = "/root/reference") -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    if pathlib.Path(reference, "src", "multiply.cpp").exists():
        subprocess.run(
            ["make", "-s", "-C", str(ROOT / "oracle"), "ref", f"REF={reference}"], check=True
        )


def build_facade_check(reference: str = "/root/reference") -> None:
    """Test program linking the C++ facade into the reference's own API
    (oracle/_ref/facade_check); needs the reference tree and the built library."""
    if pathlib.Path(reference, "include", "wolfgang", "device_pauli_sum.hpp").exists():
        subprocess.run(
            ["make", "-s", "-C", str(ROOT / "oracle"), "facade", f"REF={reference}"], check=True
        )


def main() -> int:
    build_oracle()
    lib = build_library()
    build_facade_check()
    print(f"built {lib}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
