"""ctypes binding of libpsw.so, the C ABI declared in include/psw.h.

The library is loaded from the package's own build directory
(plane-sweep-cuda_amd/lib/libpsw.so).  There is no fallback: if the library is
missing, importing this module raises, so no result can silently come from
anywhere but the HIP kernels.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.abspath(os.path.join(_HERE, "..", ".."))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libpsw.so")
INCLUDE_PATH = os.path.abspath(os.path.join(PKG_ROOT, "..", "include", "psw.h"))

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libpsw.so not found at {LIB_PATH}; build it first "
        "(python -c 'import __graft_entry__ as g; g.build()' or make -C plane-sweep-cuda_amd)")

lib = ctypes.CDLL(LIB_PATH)

c_int, c_float, c_double, c_size_t = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_size_t
c_void_p, c_char_p, c_uint = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint
P = ctypes.POINTER
c_float_p = P(c_float)
c_int64 = ctypes.c_int64


class Dim3(ctypes.Structure):
    _fields_ = [("x", c_uint), ("y", c_uint), ("z", c_uint)]


class TgvParams(ctypes.Structure):
    _fields_ = [("lambda_", c_double), ("alpha0", c_double), ("alpha1", c_double), ("tau", c_double),
                ("sigma", c_double), ("beta", c_double), ("gamma", c_double)]


class FusionGrid(ctypes.Structure):
    _fields_ = [("width", c_int), ("height", c_int), ("depth", c_int), ("volume", c_float * 6),
                ("u", c_void_p), ("v", c_void_p), ("p", c_void_p), ("hist", c_void_p)]


PSW_SCENE_MAX_BOXES = 32


class SweepFrame(ctypes.Structure):
    _fields_ = [("d_ref", c_void_p), ("d_src", c_void_p), ("width", c_int), ("height", c_int), ("nsrc", c_int),
                ("d_H", c_void_p), ("nplanes", c_int), ("d_argmin", c_void_p), ("d_cost_min", c_void_p),
                ("d_cost_second", c_void_p)]


class PswRoom(ctypes.Structure):
    _fields_ = [("lo", c_double * 3), ("hi", c_double * 3), ("nboxes", c_int),
                ("box_lo", (c_double * 3) * PSW_SCENE_MAX_BOXES), ("box_hi", (c_double * 3) * PSW_SCENE_MAX_BOXES),
                ("octave_salt", ctypes.c_uint32 * 4)]


PSW_OK = 0
PSW_ERR_ARG, PSW_ERR_LAUNCH, PSW_ERR_ALLOC, PSW_ERR_DEVICE, PSW_ERR_COPY, PSW_ERR_UNSUPPORTED = 1, 2, 3, 4, 5, 6
PSW_MAX_SOURCES = 16

D = c_void_p  # device pointer
H = c_void_p  # host pointer (numpy .ctypes.data)

_SIGS = {
    "psw_status_string": (c_char_p, [c_int]),
    "psw_last_error": (c_char_p, []),
    "psw_device_count": (c_int, [P(c_int)]),
    "psw_set_device": (c_int, [c_int]),
    "psw_device_name": (c_int, [c_char_p, c_size_t]),
    "psw_malloc": (c_int, [P(c_void_p), c_size_t]),
    "psw_free": (c_int, [D]),
    "psw_host_alloc": (c_int, [P(c_void_p), c_size_t]),
    "psw_host_free": (c_int, [H]),
    "psw_memcpy_htod": (c_int, [D, H, c_size_t, c_void_p]),
    "psw_memcpy_dtoh": (c_int, [H, D, c_size_t, c_void_p]),
    "psw_memcpy_dtod": (c_int, [D, D, c_size_t, c_void_p]),
    "psw_fill_f32": (c_int, [D, c_float, c_size_t, c_void_p]),
    "psw_fill_u32": (c_int, [D, ctypes.c_uint32, c_size_t, c_void_p]),
    "psw_copy_f32": (c_int, [D, D, c_size_t, c_void_p]),
    "psw_stream_create": (c_int, [P(c_void_p)]),
    "psw_stream_destroy": (c_int, [c_void_p]),
    "psw_stream_synchronize": (c_int, [c_void_p]),
    "psw_device_synchronize": (c_int, []),
    "psw_event_create": (c_int, [P(c_void_p)]),
    "psw_event_destroy": (c_int, [c_void_p]),
    "psw_event_record": (c_int, [c_void_p, c_void_p]),
    "psw_event_synchronize": (c_int, [c_void_p]),
    "psw_event_elapsed_ms": (c_int, [P(c_float), c_void_p, c_void_p]),
    # host helpers
    "psw_reference_planes": (c_int, [c_float, c_float, c_uint, H, c_int, P(c_int)]),
    "psw_mat3_inverse": (c_int, [H, H]),
    "psw_relative_matrices": (c_int, [H, H, H, H, c_int, H, H]),
    "psw_sweep_homographies_reference": (c_int, [H, H, c_int, c_int, H, H, H, H, H, c_int]),
    # reference launchers
    "psw_transform_indexes": (c_int, [D, D, H, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_bilinear_interpolation": (c_int, [D, D, D, D, c_int, c_int, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_calcNCC": (c_int, [D, D, D, D, D, D, c_float, c_float, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_update_arrays": (c_int, [D, D, D, c_float, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_sum_depthmap_NCC": (c_int, [D, D, D, D, c_float, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_calculate_STD": (c_int, [D, D, D, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_set_value": (c_int, [D, c_float, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_element_multiply": (c_int, [D, D, D, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_element_rdivide": (c_int, [D, D, D, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_windowed_mean_row": (c_int, [D, D, c_uint, c_int, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_windowed_mean_column": (c_int, [D, D, c_uint, c_int, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_element_scale": (c_int, [D, c_float, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_element_add": (c_int, [D, c_float, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_set_QNAN_value": (c_int, [D, c_float, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_denoising_TVL1_calculateP": (c_int, [D, D, D, c_float, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_denoising_TVL1_calculateP_tensor_weighed": (c_int, [D, D, D, D, D, D, D, c_float, c_int, c_int, Dim3, Dim3,
                                                             c_void_p]),
    "psw_denoising_TVL1_update": (c_int, [D, D, D, D, D, c_float, c_float, c_float, c_float, c_int, c_int, Dim3, Dim3,
                                          c_void_p]),
    "psw_denoising_TVL1_update_tensor_weighed": (c_int, [D, D, D, D, D, D, D, D, D, c_float, c_float, c_float, c_float,
                                                         c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_compute3D": (c_int, [D, D, D, H, H, H, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_TGV2_updateP": (c_int, [D, D, D, D, D, c_float, c_float, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_TGV2_updateQ": (c_int, [D, D, D, D, D, D, c_float, c_float, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_TGV2_updateR": (c_int, [D, D, D, D, D, D, c_float, c_float, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_TGV2_updateU": (c_int, [D, D, D, D, D, D, D, D, D, D, D, D, D, c_float, c_float, c_float, c_float, c_int,
                                 c_int, Dim3, Dim3, c_void_p]),
    "psw_TGV2_transform_coordinates": (c_int, [D, D, D, D, D, D, H, H, H, H, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_subtract": (c_int, [D, D, D, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_TGV2_calculate_coordinate_derivatives": (c_int, [D, D, D, H, H, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_TGV2_calculate_derivativeF": (c_int, [D, D, D, D, D, D, D, D, c_float, c_float, c_int, c_int, Dim3, Dim3,
                                               c_void_p]),
    "psw_TGV2_calculate_Iu": (c_int, [D, D, D, D, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_Anisotropic_diffusion_tensor": (c_int, [D, D, D, D, D, c_float, c_float, c_int, c_int, Dim3, Dim3,
                                                 c_void_p]),
    "psw_TGV2_updateU_tensor_weighed": (c_int, [D] * 17 + [c_float, c_float, c_float, c_float, c_int, c_int, Dim3,
                                                           Dim3, c_void_p]),
    "psw_TGV2_updateP_tensor_weighed": (c_int, [D] * 9 + [c_float, c_float, c_int, c_int, Dim3, Dim3, c_void_p]),
    "psw_TGV2_updateU_sparseDepth": (c_int, [D] * 14 + [c_float, c_float, c_float, c_float, c_int, c_int, Dim3, Dim3,
                                                        c_void_p]),
    "psw_TGV2_updateU_sparseDepthTensor": (c_int, [D] * 18 + [c_float, c_float, c_float, c_float, c_int, c_int, Dim3,
                                                              Dim3, c_void_p]),
    "psw_calculateWeights_sparseDepth": (c_int, [D, D, c_int, c_int, Dim3, Dim3, c_void_p]),
    # fused sweep
    "psw_sweep_homographies": (c_int, [D, D, c_int, c_int, H, H, H, H, H, c_int, c_void_p]),
    "psw_sweep_fused_reference": (c_int, [D, D, c_int, c_int, c_int, D, D, c_int, c_int, c_float, c_float, c_float, D,
                                          c_void_p]),
    "psw_sweep_fused_aggregate": (c_int, [D, D, c_int, c_int, c_int, D, c_int, c_int, c_float, c_int, c_int, D, D, D,
                                          c_void_p]),
    "psw_sweep_fused_aggregate_noalloc": (c_int, [D, D, c_int, c_int, c_int, D, c_int, c_int, c_float, c_int,
                                                  c_int, c_int, D, D, D, c_void_p]),
    "psw_sweep_argmin_depth": (c_int, [D, D, D, c_int, c_float, c_void_p]),
    # TGV2
    "psw_tgv_create": (c_int, [P(c_void_p), c_int, c_int, c_int]),
    "psw_tgv_destroy": (c_int, [c_void_p]),
    "psw_tgv_state_bytes": (c_size_t, [c_int, c_int, c_int]),
    "psw_tgv_set_images": (c_int, [c_void_p, D, D, c_double, c_double, c_void_p]),
    "psw_tgv_set_u": (c_int, [c_void_p, D, c_float, c_void_p]),
    "psw_tgv_set_primal": (c_int, [c_void_p, D, c_float, D, c_void_p]),
    "psw_tgv_set_duals": (c_int, [c_void_p, D, D, D, c_void_p]),
    "psw_tgv_set_energy_mask": (c_int, [c_void_p, c_double, P(TgvParams), P(ctypes.c_int64), c_void_p]),
    "psw_tgv_iterate_batch": (c_int, [P(c_void_p), c_int, c_int, P(TgvParams), c_int, c_void_p]),
    "psw_tgv_solve_batch": (c_int, [P(c_void_p), c_int, P(TgvParams), P(c_double), c_int, c_int, c_double, c_int, c_int,
                                    P(c_int), P(c_double), c_void_p]),
    "psw_sweep_fused_aggregate_batch": (c_int, [P(SweepFrame), c_int, c_int, c_float, c_int, c_int, c_int, c_int,
                                                c_void_p]),
    "psw_tgv_linearise": (c_int, [c_void_p, H, H, H, c_void_p]),
    "psw_tgv_iterate": (c_int, [c_void_p, c_int, P(TgvParams), c_void_p]),
    "psw_tgv_energy": (c_int, [c_void_p, P(TgvParams), P(c_double), c_void_p]),
    "psw_tgv_solve": (c_int, [c_void_p, c_int, c_int, c_double, c_double, P(TgvParams), P(c_int), P(c_double),
                              c_void_p]),
    "psw_tgv_get_u": (c_int, [c_void_p, D, c_void_p]),
    "psw_tgv_get_state": (c_int, [c_void_p, D, c_void_p]),
    "psw_tgv_get_energy_mask": (c_int, [c_void_p, D, c_void_p]),
    "psw_tgv_run": (c_int, [c_void_p, H, H, H, c_int, c_int, P(TgvParams), D, D, c_void_p]),
    "psw_tgv_sparse_workspace_bytes": (c_size_t, [c_int, c_int]),
    "psw_tgv_sparse_denoise": (c_int, [D, D, D, c_int, c_int, c_float, c_int] + [c_double] * 7 + [D, c_size_t, D,
                                                                                              c_void_p]),
    # fusion
    "psw_fusion_alloc": (c_int, [P(FusionGrid), c_int, c_int, c_int, H]),
    "psw_fusion_free": (c_int, [P(FusionGrid)]),
    "psw_fusion_clear": (c_int, [P(FusionGrid), c_void_p]),
    "psw_fusion_update_histogram": (c_int, [P(FusionGrid), D, c_int, c_int, H, H, H, c_float, c_void_p]),
    "psw_fusion_update_histogram_batch": (c_int, [P(FusionGrid), D, c_int, c_int, c_int, H, H, H, c_float,
                                                  c_void_p]),
    "psw_fusion_update_u": (c_int, [P(FusionGrid), c_double, c_double, c_void_p]),
    "psw_fusion_update_p": (c_int, [P(FusionGrid), c_double, c_void_p]),
    "psw_fusion_update_iteration": (c_int, [P(FusionGrid), D, c_int, c_int, H, H, H, c_float, c_double, c_double,
                                            c_double, c_void_p]),
    "psw_fusion_solve": (c_int, [P(FusionGrid), c_int, c_double, c_double, c_double, c_void_p]),
    "psw_fusion_energy": (c_int, [P(FusionGrid), c_double, P(c_double), c_void_p]),
    "psw_fusion_surface_points": (c_int, [P(FusionGrid), D, ctypes.c_int64, P(ctypes.c_int64), c_void_p]),
    "psw_fusion_surface_count": (c_int, [P(FusionGrid), P(c_int64), c_void_p]),
    # TVL1
    "psw_scene_render": (c_int, [P(PswRoom), H, H, H, c_int, c_int, c_double, ctypes.c_uint64, c_float, D, D,
                                 c_void_p]),
    # learned warm start
    "psw_unet_param_count": (c_size_t, []),
    "psw_unet_create": (c_int, [P(c_void_p), ctypes.c_int64]),
    "psw_unet_destroy": (c_int, [c_void_p]),
    "psw_unet_init": (c_int, [c_void_p, ctypes.c_uint64]),
    "psw_unet_set_params": (c_int, [c_void_p, H, c_size_t]),
    "psw_unet_get_params": (c_int, [c_void_p, H, c_size_t]),
    "psw_unet_get_grads": (c_int, [c_void_p, H, c_size_t]),
    "psw_unet_features": (c_int, [c_void_p, c_int, c_int, c_int, D, D, D, D, D, c_int, c_float, c_void_p]),
    "psw_unet_set_input": (c_int, [c_void_p, c_int, c_int, c_int, D, D, c_void_p]),
    "psw_unet_targets": (c_int, [c_void_p, c_int, c_int, c_int, D, c_void_p]),
    "psw_unet_set_target": (c_int, [c_void_p, c_int, c_int, c_int, D, c_void_p]),
    "psw_unet_forward": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p]),
    "psw_unet_output": (c_int, [c_void_p, c_int, c_int, c_int, D, c_void_p]),
    "psw_unet_warm_start": (c_int, [c_void_p, c_int, c_int, c_int, D, D, D, D, c_void_p]),
    "psw_unet_loss_grad": (c_int, [c_void_p, c_int, c_int, c_int, H, P(c_double), c_void_p]),
    "psw_unet_adam_step": (c_int, [c_void_p, c_double, c_double, c_double, c_double, c_int, c_void_p]),
    "psw_unet_flops": (c_int, [c_int, c_int, c_int, P(c_double)]),
    "psw_conv3x3_bf16": (c_int, [D, c_int, D, c_int, D, D, c_int, c_int, c_int, c_int, c_int, D, c_void_p]),
    "psw_conv3x3_bf16_dgrad": (c_int, [D, c_int, D, c_int, c_int, c_int, c_int, D, c_void_p]),
    "psw_conv3x3_bf16_wgrad": (c_int, [D, c_int, D, c_int, D, c_int, c_int, c_int, c_int, D, c_void_p]),
    "psw_tvl1_denoise": (c_int, [D, D, c_int, c_int, c_float, c_float, c_int, c_double, c_double, c_double, c_double,
                                 c_double, c_double, D, D, c_void_p]),
}

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class PswError(RuntimeError):
    """A libpsw entry point returned a non-zero status."""

    def __init__(self, func, status):
        msg = lib.psw_last_error().decode(errors="replace")
        name = lib.psw_status_string(status).decode()
        super().__init__(f"{func}: {name} ({status}): {msg}")
        self.status = status


def check(func_name, status):
    if status != PSW_OK:
        raise PswError(func_name, status)
    return status


def call(func_name, *args):
    """Call psw_<func_name>(*args) and raise PswError on a non-zero status."""
    return check(func_name, getattr(lib, func_name)(*args))


ZERO_DIM = Dim3(0, 0, 0)


def dim3(x=0, y=0, z=0):
    return Dim3(x, y, z)
