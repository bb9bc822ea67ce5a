"""The header-only C++ HipTensor (include/jet_hip/HipTensor.hpp).

HipTensor is the C++ drop-in for the reference's CudaTensor
(include/jet/CudaTensor.hpp); tests/native/hip_tensor_check.cpp runs the
reference's CudaTensor test cases and literal values
(test/Test_CudaTensor.cpp:21-678) against it through the C ABI.

CPU: the header compiles with a plain host compiler (g++, no HIP headers) and
the check library links and loads.  GPU: every case passes.
"""
import ctypes
import os
import subprocess

import pytest

from jet_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "native", "_build", "libhip_tensor_check.so")


def _lib():
    if not os.path.exists(LIB):
        pytest.skip("tests/native/_build/libhip_tensor_check.so not built (run build())")
    _native.lib()  # the library HipTensor links against, loaded first
    lib = ctypes.CDLL(LIB)
    lib.hip_tensor_check.restype = ctypes.c_int
    lib.hip_tensor_check.argtypes = [ctypes.c_char_p, ctypes.c_int]
    return lib


def test_header_compiles_without_hip(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "jet_hip/HipTensor.hpp"\n'
                   "template class Jet::HipTensor<std::complex<float>>;\n"
                   "template class Jet::HipTensor<std::complex<double>, 1>;\n")
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                    "-I" + os.path.join(ROOT, "include"), str(src)], check=True)


def test_check_library_loads():
    assert hasattr(_lib(), "hip_tensor_check")


@pytest.mark.gpu
def test_hip_tensor_reference_cases():
    lib = _lib()
    buf = ctypes.create_string_buffer(1024)
    rc = lib.hip_tensor_check(buf, len(buf))
    msg = buf.value.decode()
    print(msg, flush=True)
    assert rc == 0, msg


REF_NET = os.path.join(ROOT, "oracle", "_ref", "libref_network.so")


def _ref_network_lib():
    if not os.path.exists(REF_NET):
        pytest.skip("oracle/_ref/libref_network.so not built (needs the reference tree)")
    _native.lib()
    lib = ctypes.CDLL(REF_NET)
    lib.jr_network_hiptensor_check.restype = ctypes.c_int
    lib.jr_network_hiptensor_check.argtypes = [ctypes.c_char_p, ctypes.c_int]
    return lib


def test_reference_network_templates_library_loads():
    assert hasattr(_ref_network_lib(), "jr_network_hiptensor_check")


@pytest.mark.gpu
def test_reference_network_templates_over_hip_tensor():
    """The reference's own TensorNetwork / PathInfo templates
    (TensorNetwork.hpp, PathInfo.hpp) instantiated with HipTensor reproduce
    the values the reference's tests expect of TensorNetwork<CudaTensor>."""
    lib = _ref_network_lib()
    buf = ctypes.create_string_buffer(1024)
    rc = lib.jr_network_hiptensor_check(buf, len(buf))
    msg = buf.value.decode()
    print(msg, flush=True)
    assert rc == 0, msg
