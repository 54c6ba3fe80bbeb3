"""The fused TGV2 kernels project the duals with their own correctly rounded
square root and division (plane-sweep-cuda_amd/csrc/tgv_point.h:
sqrt_rn_gt1, rcp_nr, div_nr) instead of the general sqrtf and '/'.  This
checks them bit for bit against sqrtf and '/' on the GPU, over 16.7 M
hashed cases per seed: roots over every binade above 1, quotients with
numerators down to the denormals, and whole dual updates against the
reference formula x / fmaxf(1, sqrt(s)) (TGV2_kernels.cu,
TGV2_updateP_tensor_weighed / TGV2_updateQ).  The kernel and its driver
are tests/cpp/point_check.hip."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "tests", "cpp", "_build", "libpoint_check.so")


def test_point_check_library_exports():
    lib = ctypes.CDLL(LIB)
    assert hasattr(lib, "point_check")


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2])
def test_projection_arithmetic_bit_exact(gpu, seed):
    lib = ctypes.CDLL(LIB)
    lib.point_check.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_ulonglong)]
    lib.point_check.restype = ctypes.c_int
    n = 1 << 24
    out = (ctypes.c_ulonglong * 5)()
    print(f"[case] point_check n={n} seed={seed}", flush=True)
    assert lib.point_check(n, seed, out) == 0
    bad_root, bad_div, fast, bad_pq, proj = (int(v) for v in out)
    print(f"root mismatches {bad_root}, quotient mismatches {bad_div} of {fast} on the fast path, "
          f"dual-update mismatches {bad_pq} ({proj} projected)")
    assert bad_root == 0 and bad_div == 0 and bad_pq == 0
    # the fast quotient path ran (|x| >= 2^-100 and s < 2^40: about 1 case in 8),
    # and dual updates with and without a projection were compared
    assert fast > n // 16 and n // 8 < proj < n - n // 8
