"""The C ABI library loads and exports every function include/psw.h declares
(no GPU needed: nothing is called that touches a device)."""
import ctypes
import os
import re
import subprocess

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "psw.h")
LIB = os.path.join(ROOT, "plane-sweep-cuda_amd", "lib", "libpsw.so")
CXXLIB = os.path.join(ROOT, "plane-sweep-cuda_amd", "lib", "libpsw_cxx.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(psw_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_launchers():
    names = declared_functions()
    # every kernel launcher of src/inc/kernels.cu.h used by the drivers
    for k in ["transform_indexes", "bilinear_interpolation", "calcNCC", "update_arrays", "sum_depthmap_NCC",
              "calculate_STD", "set_value", "element_multiply", "element_rdivide", "windowed_mean_row",
              "windowed_mean_column", "element_scale", "element_add", "set_QNAN_value",
              "denoising_TVL1_calculateP_tensor_weighed", "denoising_TVL1_update", "TGV2_updateP_tensor_weighed",
              "TGV2_updateQ", "TGV2_updateR", "TGV2_updateU_tensor_weighed", "TGV2_transform_coordinates",
              "subtract", "TGV2_calculate_coordinate_derivatives", "TGV2_calculate_derivativeF", "TGV2_calculate_Iu",
              "Anisotropic_diffusion_tensor", "fusion_update_histogram", "fusion_update_u", "fusion_update_p",
              "fusion_update_iteration"]:
        assert "psw_" + k in names, k
    assert len(names) > 80


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libpsw.so not built"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_exports_with_nm():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    assert set(declared_functions()) <= exported


def test_status_strings_without_device():
    lib = ctypes.CDLL(LIB)
    lib.psw_status_string.restype = ctypes.c_char_p
    assert lib.psw_status_string(0) == b"ok"
    assert lib.psw_status_string(1) == b"invalid argument"


def test_host_helpers_run_without_device():
    """Host-only entry points (plane list, relative pose) need no GPU."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "plane-sweep-cuda_amd", "python"))
    import numpy as np
    import oracle
    from psw_amd import stages
    for zn, zf, n in [(0.1, 1.0, 200), (0.5, 5.0, 64), (0.3, 8.0, 256), (0.1, 5.0, 250), (1.0, 2.0, 3)]:
        np.testing.assert_array_equal(stages.reference_planes(zn, zf, n), oracle.reference_planes(zn, zf, n))
    rng = np.random.default_rng(0)
    from psw_amd import scene
    for alt in (0, 1):
        R1, R2 = scene.rotation(rng.normal(size=3), 0.3), scene.rotation(rng.normal(size=3), 0.2)
        t1, t2 = rng.normal(size=3).astype(np.float32), rng.normal(size=3).astype(np.float32)
        a = stages.relative_matrices(R1, t1, R2, t2, alt)
        b = oracle.relative_matrices(R1, t1, R2, t2, alt)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
    K = scene.one_based(scene.intrinsics())
    Hs = stages.homographies_reference(np.array([0.7, 2.0], np.float32), K, np.eye(3), np.zeros(3), R2[None], t2[None])
    for p, d in enumerate([0.7, 2.0]):
        np.testing.assert_array_equal(Hs[0, p].reshape(3, 3),
                                      oracle.homography_ref(K, np.eye(3), np.zeros(3), R2, t2, np.float32(d)))
