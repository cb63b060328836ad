"""The C++ facade (wolfgang_amd/facade/wolfgang_hip_amd.cpp) as a drop-in for
the reference's HIP backend.

* It compiles against the reference's own public headers
  (include/wolfgang/device_pauli_sum.hpp, device_commutation_matrix.hpp) with
  -Werror, and defines every member those headers declare.
* oracle/_ref/facade_check (tests/facade_check.cpp linked with the facade,
  libwolfgang_amd.so and the reference's CPU sources) runs the reference's C++
  API: without a device it checks the status members and errors; on the GPU it
  checks every member against the reference's own CPU PauliSum (bit-exact
  algebra and commutation, expectations within 1e-12 * sum|c|).
"""

import os
import pathlib
import re
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
FACADE = ROOT / "wolfgang_amd" / "facade" / "wolfgang_hip_amd.cpp"
CHECK = ROOT / "oracle" / "_ref" / "facade_check"
REF_INCLUDE = pathlib.Path("/root/reference/include")

# Members declared by reference include/wolfgang/device_pauli_sum.hpp:93-175 and
# include/wolfgang/device_commutation_matrix.hpp:16-66.
PAULI_MEMBERS = [
    "DevicePauliSum()", "~DevicePauliSum()", "DevicePauliSum(wolfgang::DevicePauliSum&&)",
    "operator=(wolfgang::DevicePauliSum&&)", "DevicePauliSum(std::unique_ptr<",
    "from_host(wolfgang::PauliSum const&, int)",
    "from_host(wolfgang::PauliSum const&, wolfgang::AcceleratorBackend, int)",
    "cuda_status()", "cuda_available()", "hip_status()", "hip_available()", "metal_status()",
    "metal_available()", "to_host() const", "simplify(double, double) const",
    "expectation_statevector_complex128(", "expectation_statevector_complex64(",
    "expectation_statevector_device_pointer(", "expectation_statevector_device_pointer_impl(",
    "commutes_with(", "commutes_with_into(", "commutes_with_device(", "commutes_with_device_into(",
    "matmul(", "num_qubits() const", "num_terms() const", "words() const", "device() const",
    "backend[abi:cxx11]() const",
]
MATRIX_MEMBERS = [
    "DeviceCommutationMatrix()", "~DeviceCommutationMatrix()",
    "DeviceCommutationMatrix(wolfgang::DeviceCommutationMatrix&&)",
    "operator=(wolfgang::DeviceCommutationMatrix&&)", "DeviceCommutationMatrix(std::unique_ptr<",
    "empty(unsigned long, unsigned long, int)",
    "empty(unsigned long, unsigned long, wolfgang::AcceleratorBackend, int)",
    "count_commuting() const", "count_commuting_rows() const", "count_commuting_cols() const",
    "to_host() const", "rows() const", "cols() const", "num_entries() const", "device() const",
    "backend[abi:cxx11]() const", "data_pointer_for_cuda_array_interface() const",
    "data_pointer_for_dlpack() const", "dlpack_device_type() const",
    "mutable_data_for_device_write()",
]


@pytest.mark.skipif(not REF_INCLUDE.exists(), reason="reference headers not mounted")
def test_facade_defines_every_reference_member(tmp_path):
    obj = tmp_path / "facade.o"
    cmd = ["g++", "-std=c++20", "-Wall", "-Wextra", "-Werror", "-c", f"-I{REF_INCLUDE}",
           f"-I{ROOT / 'include'}", str(FACADE), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr
    nm = subprocess.run(["nm", "-C", "--defined-only", str(obj)], capture_output=True, text=True,
                        check=True).stdout
    defined = {m.group(1) for m in re.finditer(r" T (wolfgang::.*)$", nm, re.M)}
    for cls, members in (("DevicePauliSum", PAULI_MEMBERS), ("DeviceCommutationMatrix", MATRIX_MEMBERS)):
        for member in members:
            prefix = f"wolfgang::{cls}::{member}"
            assert any(d.startswith(prefix) for d in defined), f"{prefix} is not defined"
    assert any(d.startswith("wolfgang::copy_device_commutation_matrix_from_host_for_testing(")
               for d in defined)


@pytest.mark.skipif(not CHECK.exists(), reason="facade check not built (needs the reference tree)")
@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present; the device mode covers it")
def test_facade_check_without_device():
    proc = subprocess.run([str(CHECK), "nodevice"], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "facade_check: OK" in proc.stdout


@pytest.mark.gpu
@pytest.mark.skipif(not CHECK.exists(), reason="facade check not built (needs the reference tree)")
def test_facade_check_on_device(gpu):
    print("facade_check device", flush=True)
    proc = subprocess.run([str(CHECK), "device"], capture_output=True, text=True, timeout=180)
    print(proc.stdout, flush=True)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "facade_check: OK" in proc.stdout


def test_facade_check_source_has_no_reference_path():
    # The program runs on machines without the reference tree: it names no
    # absolute reference path.
    assert "/root/reference" not in (ROOT / "tests" / "facade_check.cpp").read_text()
    assert shutil.which("g++") is not None
