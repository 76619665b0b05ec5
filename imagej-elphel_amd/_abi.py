"""ctypes binding of the C ABI in include/clt_tile_processor.h (libclt_hip.so).

The library is the product: there is no CPU fallback.  Loading fails loudly when the built
library is missing.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclt_hip.so")

CLT_OK = 0
ERRORS = {-1: "CLT_E_ARG", -2: "CLT_E_STATE", -3: "CLT_E_HIP", -4: "CLT_E_NODEVICE",
          -5: "CLT_E_NOMEM", -6: "CLT_E_UNSUPPORTED"}
DISP_LAYERS = 12
CORR_ADJ = 32
DISP_NAMES = ["int", "cm", "strength", "hor", "hor_strength", "vert", "vert_strength", "cm_weight",
              "poly", "poly_strength", "cm_raw", "strength_raw"]
TIMING_KERNELS = {"convert": 0, "imclt": 1, "corr2d": 2, "textures": 3, "combine_textures": 4, "batch": 5,
                  "corr_adjoint": 6, "geometry_adjoint": 7, "corr_lma": 8,
                  "kernel_prep": 9, "rig": 10, "tex_imclt": 11, "convert_corr2d": 12}


class clt_geom(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "width", "height", "num_cams", "num_colors", "red_row", "red_col",
        "kernels_hor", "kernels_vert", "kernel_step", "max_tasks")]


class clt_pair(ctypes.Structure):
    _fields_ = [("cam1", ctypes.c_int32), ("cam2", ctypes.c_int32),
                ("dir", ctypes.c_int32), ("scale", ctypes.c_int32)]


class clt_corr_params(ctypes.Structure):
    _fields_ = [("fat_zero", ctypes.c_float), ("corr_red", ctypes.c_float),
                ("corr_blue", ctypes.c_float), ("corr_offset", ctypes.c_float),
                ("twice_diagonal", ctypes.c_int32), ("min_corr", ctypes.c_float),
                ("strip_height", ctypes.c_int32), ("enhortho_width", ctypes.c_float),
                ("enhortho_scale", ctypes.c_float)]


class clt_rig_params(ctypes.Structure):
    _fields_ = [("use_poly", ctypes.c_int32), ("use_xy_poly", ctypes.c_int32), ("no_int_x0", ctypes.c_int32),
                ("notch_mode", ctypes.c_int32), ("lt_rad", ctypes.c_int32), ("corr_strip_notch", ctypes.c_int32),
                ("corr_notch_hwidth", ctypes.c_double), ("corr_notch_blur", ctypes.c_double),
                ("cm_max_normalization", ctypes.c_double), ("max_corr_radius", ctypes.c_double)]


BI_NAMES = ["disp_full", "disp_hor", "disp_vert", "disp_diagm", "disp_diago",
            "adisp_full", "adisp_hor", "adisp_vert", "adisp_diagm", "adisp_diago",
            "disp_cross", "disp_cross_dx", "disp_cross_dy",
            "str_full", "str_hor", "str_vert", "str_diagm", "str_diago",
            "astr_full", "astr_hor", "astr_vert", "astr_diagm", "astr_diago",
            "str_cross", "str_all", "target"]


class clt_corr_lma_params(ctypes.Structure):
    _fields_ = [("enable", ctypes.c_int32), ("mix_corr_poly", ctypes.c_int32),
                ("min_poly_strength", ctypes.c_double), ("ortho_vasw_pwr", ctypes.c_double),
                ("cnvx_hwnd_size", ctypes.c_int32), ("cnvx_add3x3", ctypes.c_int32),
                ("cnvx_weight", ctypes.c_double),
                ("lma_adjust_wm", ctypes.c_int32), ("lma_adjust_wy", ctypes.c_int32),
                ("lma_adjust_wxy", ctypes.c_int32), ("lma_adjust_ag", ctypes.c_int32),
                ("lma_half_width", ctypes.c_double), ("lma_cost_wy", ctypes.c_double),
                ("lma_cost_wxy", ctypes.c_double), ("lma_lambda_initial", ctypes.c_double),
                ("lma_lambda_scale_good", ctypes.c_double), ("lma_lambda_scale_bad", ctypes.c_double),
                ("lma_lambda_max", ctypes.c_double), ("lma_rms_diff", ctypes.c_double),
                ("lma_num_iter", ctypes.c_int32), ("fo_correct", ctypes.c_int32),
                ("fo_far", ctypes.c_int32), ("fo_ortho", ctypes.c_int32),
                ("fo_min_strength", ctypes.c_double), ("fo_min_eff", ctypes.c_double),
                ("fo_min_eff_ratio", ctypes.c_double), ("fo_max_hwidth", ctypes.c_double),
                ("fo_min_diff", ctypes.c_double), ("fo_overcorrection", ctypes.c_double),
                ("fo_lim_overcorr", ctypes.c_double)]


class clt_tex_params(ctypes.Structure):
    _fields_ = [("min_shot", ctypes.c_float), ("scale_shot", ctypes.c_float),
                ("diff_sigma", ctypes.c_float), ("dust_remove", ctypes.c_int32),
                ("corr_magic_scale", ctypes.c_float), ("keep_weights", ctypes.c_int32),
                ("port_offsets", (ctypes.c_float * 2) * 16)]


class clt_port_geometry(ctypes.Structure):
    _fields_ = [("num_sensors", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("focal_length", ctypes.c_double), ("pixel_size", ctypes.c_double),
                ("distortion_radius", ctypes.c_double), ("distortion", ctypes.c_double * 7),
                ("rxy", (ctypes.c_double * 2) * 4), ("pxy0", (ctypes.c_double * 2) * 4),
                ("roll_deg", ctypes.c_double * 4), ("extrinsic", ctypes.c_double * 13)]


class clt_rig_offset(ctypes.Structure):
    _fields_ = [("baseline", ctypes.c_double), ("aux_angle", ctypes.c_double), ("aux_z", ctypes.c_double),
                ("aux_azimuth", ctypes.c_double), ("aux_tilt", ctypes.c_double), ("aux_roll", ctypes.c_double),
                ("aux_zoom", ctypes.c_double), ("main_disparity_radius", ctypes.c_double),
                ("aux_disparity_radius", ctypes.c_double)]


class clt_lma_params(ctypes.Structure):
    _fields_ = [("lambda_", ctypes.c_double), ("lambda_scale_good", ctypes.c_double),
                ("lambda_scale_bad", ctypes.c_double), ("lambda_max", ctypes.c_double),
                ("rms_diff", ctypes.c_double), ("num_iter", ctypes.c_int32), ("adjust_disparity", ctypes.c_int32),
                ("coarse_iter", ctypes.c_int32), ("reserved", ctypes.c_int32), ("coarse_radius", ctypes.c_double),
                ("coarse_strength_floor", ctypes.c_double)]


_P = ctypes.c_void_p
_H = ctypes.c_void_p
_I = ctypes.c_int32

# name -> (restype, argtypes); every entry point declared in include/clt_tile_processor.h
SIGNATURES = {
    "clt_create": (_I, [ctypes.POINTER(clt_geom), ctypes.POINTER(_H)]),
    "clt_destroy": (_I, [_H]),
    "clt_last_error": (ctypes.c_char_p, [_H]),
    "clt_device_count": (_I, []),
    "clt_set_kernels": (_I, [_H, _I, _P, _P]),
    "clt_clear_kernels": (_I, [_H]),
    "clt_prepare_kernels": (_I, [_H, _I, _P, _I, _P, _I]),
    "clt_get_kernels": (_I, [_H, _I, _P, _P]),
    "clt_kernels_to_flat": (_I, [_P, _P, _I, _I, _I, _P]),
    "clt_kernels_from_flat": (_I, [_P, _I, _I, _I, _P, _P]),
    "clt_set_image": (_I, [_H, _I, _P, _I]),
    "clt_set_tasks": (_I, [_H, _P, _I]),
    "clt_set_lpf_rbg": (_I, [_H, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "clt_set_corr_sigma": (_I, [_H, ctypes.c_float]),
    "clt_set_pairs": (_I, [_H, _P, _I]),
    "clt_set_corr_params": (_I, [_H, ctypes.POINTER(clt_corr_params)]),
    "clt_set_tex_params": (_I, [_H, ctypes.POINTER(clt_tex_params)]),
    "clt_set_corr_lma_params": (_I, [_H, ctypes.POINTER(clt_corr_lma_params)]),
    "clt_lpf_coeffs": (_I, [ctypes.c_double, _P]),
    "clt_default_corr_params": (None, [ctypes.POINTER(clt_corr_params)]),
    "clt_default_tex_params": (None, [ctypes.POINTER(clt_tex_params), _I]),
    "clt_default_corr_lma_params": (None, [ctypes.POINTER(clt_corr_lma_params)]),
    "clt_default_rig_params": (None, [ctypes.POINTER(clt_rig_params)]),
    "clt_set_rig_params": (_I, [_H, ctypes.POINTER(clt_rig_params)]),
    "clt_exec_corr2d_rig": (_I, [_H, _P]),
    "clt_get_corr2d_rig": (_I, [_H, _P, _P, _P, _P, _P, _P]),
    "clt_exec_textures_rig": (_I, [_H]),
    "clt_get_textures_rig": (_I, [_H, _P, _P, _P, _P]),
    "clt_default_quad_pairs": (_I, [_P]),
    "clt_exec_convert": (_I, [_H, _I]),
    "clt_exec_imclt": (_I, [_H, _I]),
    "clt_exec_corr2d": (_I, [_H]),
    "clt_exec_convert_corr2d": (_I, [_H, _I]),
    "clt_exec_textures": (_I, [_H, _P]),
    "clt_exec_textures_imclt": (_I, [_H, _P, _I]),
    "clt_exec_combine_textures": (_I, [_H]),
    "clt_exec_combine_rgba_tiles": (_I, [_H, _I, _I]),
    "clt_get_rbg": (_I, [_H, _I, _P]),
    "clt_get_clt": (_I, [_H, _I, _P]),
    "clt_get_corr2d": (_I, [_H, _P, _P, _P]),
    "clt_get_corr_offsets": (_I, [_H, ctypes.c_double, _P]),
    "clt_get_textures": (_I, [_H, _P, _P]),
    "clt_get_combined_textures": (_I, [_H, _P]),
    "clt_set_async": (_I, [_H, _I]),
    "clt_sync": (_I, [_H]),
    "clt_set_timing": (_I, [_H, _I]),
    "clt_get_timing": (_I, [_H, _I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I)]),
    "clt_reset_timing": (_I, [_H]),
    "clt_default_port_geometry": (None, [ctypes.POINTER(clt_port_geometry), _I, _I]),
    "clt_set_port_geometry": (_I, [_H, ctypes.POINTER(clt_port_geometry)]),
    "clt_set_port_geometry_rig": (_I, [_H, ctypes.POINTER(clt_port_geometry), ctypes.POINTER(clt_port_geometry),
                                       ctypes.POINTER(clt_rig_offset)]),
    "clt_set_tasks_geometry": (_I, [_H, _P, _P, _P, _I]),
    "clt_get_tasks": (_I, [_H, _P]),
    "clt_get_port_derivatives": (_I, [_H, _P]),
    "clt_create_batch": (_I, [ctypes.POINTER(clt_geom), _I, ctypes.POINTER(_H)]),
    "clt_set_image_frame": (_I, [_H, _I, _I, _P, _I]),
    "clt_set_tasks_frame": (_I, [_H, _I, _P, _I]),
    "clt_exec_batch": (_I, [_H, _I]),
    "clt_autotune": (_I, [_H]),
    "clt_get_tune": (_I, [_H, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "clt_set_batch_chunks": (_I, [_H, _I]),
    "clt_get_batch_chunks": (_I, [_H, ctypes.POINTER(_I)]),
    "clt_get_tune_all": (_I, [_H, ctypes.POINTER(_I), _I]),
    "clt_set_tune_all": (_I, [_H, ctypes.POINTER(_I), _I]),
    "clt_get_corr2d_frame": (_I, [_H, _I, _P, _P, _P]),
    "clt_get_textures_frame": (_I, [_H, _I, _P, _P]),
    "clt_get_rbg_frame": (_I, [_H, _I, _I, _P]),
    "clt_exec_corr_adjoint": (_I, [_H]),
    "clt_get_corr_adjoint": (_I, [_H, _P]),
    "clt_exec_geometry_adjoint": (_I, [_H, _P, _P, _P]),
    "clt_exec_geometry_adjoint_noalloc": (_I, [_H, _P, _P, _P]),
    "clt_default_lma_params": (None, [ctypes.POINTER(clt_lma_params)]),
    "clt_lma_extrinsics": (_I, [_H, ctypes.POINTER(clt_lma_params), _P, _P, _P, ctypes.POINTER(_I)]),
}

_lib = None


def load():
    """Load libclt_hip.so and declare every exported function; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libclt_hip.so is not built (%s); run __graft_entry__.build() or "
                          "python imagej-elphel_amd/build.py" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class CltError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code
