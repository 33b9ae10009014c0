"""Division by the grid extents through a multiply-high (the FP32 push
decodes a cell id into padded indices that way): exact against integer
division for every divisor a grid can have and numerators up to 2^31."""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "fastdiv_check.cpp")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("fd") / "libfd.so")
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-std=c++17", SRC, "-o", so], check=True)
    L = C.CDLL(so)
    L.fastdiv_mismatches.restype = C.c_longlong
    L.fastdiv_mismatches.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    return L


def test_all_numerators_below_2_26_small_divisors(lib):
    # every cell id of the FP32 cell word (26 bits) for the extents 1..20
    for d in range(1, 21):
        assert lib.fastdiv_mismatches(d, 1 << 26, 1) == 0, d


@pytest.mark.parametrize("d", [66, 67, 130, 131, 514, 515, 1000, 4097, 65535, 65536, 65537, 1 << 20, (1 << 26) - 3])
def test_numerators_up_to_2_31(lib, d):
    assert lib.fastdiv_mismatches(d, (1 << 31) - 1, 7919) == 0
    assert lib.fastdiv_mismatches(d, 1 << 26, 1) == 0
