"""CPU oracle of the CLT tile processor -- TEST INFRASTRUCTURE ONLY.

ctypes front end of ``clt_oracle.c``, the C restatement of the reference's Java CPU path
(DttRad2 / ImageDtt / Correlation2d).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the checker or the timed
CPU baseline; the product path never does.

Parity pinning of each oracle function (golden tables held by the reference, closed forms,
invariants) is documented and tested in ``tests/test_oracle.py``.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


class Geom(ctypes.Structure):
    """Mirror of ``clt_geom`` in include/clt_tile_processor.h."""
    _fields_ = [(n, ctypes.c_int32) for n in (
        "width", "height", "num_cams", "num_colors", "red_row", "red_col",
        "kernels_hor", "kernels_vert", "kernel_step", "max_tasks")]


class Pair(ctypes.Structure):
    _fields_ = [("cam1", ctypes.c_int32), ("cam2", ctypes.c_int32),
                ("dir", ctypes.c_int32), ("scale", ctypes.c_int32)]


class CorrParams(ctypes.Structure):
    _fields_ = [("fat_zero", ctypes.c_float), ("corr_red", ctypes.c_float),
                ("corr_blue", ctypes.c_float), ("corr_offset", ctypes.c_float),
                ("twice_diagonal", ctypes.c_int32), ("min_corr", ctypes.c_float),
                ("strip_height", ctypes.c_int32), ("enhortho_width", ctypes.c_float),
                ("enhortho_scale", ctypes.c_float)]


class CorrLmaParams(ctypes.Structure):
    """Mirror of ``clt_corr_lma_params`` (ImageDttParameters IDP:32-110)."""
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


class RigParams(ctypes.Structure):
    """Mirror of ``clt_rig_params`` (BiQuadParameters BQP:32-83, IDP:81-112, ECP:2566)."""
    _fields_ = [("use_poly", ctypes.c_int32), ("use_xy_poly", ctypes.c_int32), ("no_int_x0", ctypes.c_int32),
                ("notch_mode", ctypes.c_int32), ("lt_rad", ctypes.c_int32), ("corr_strip_notch", ctypes.c_int32),
                ("corr_notch_hwidth", ctypes.c_double), ("corr_notch_blur", ctypes.c_double),
                ("cm_max_normalization", ctypes.c_double), ("max_corr_radius", ctypes.c_double)]


class RigOffset(ctypes.Structure):
    """Mirror of ``clt_rig_offset`` (GeometryCorrection.RigOffset GC:256-290)."""
    _fields_ = [("baseline", ctypes.c_double), ("aux_angle", ctypes.c_double), ("aux_z", ctypes.c_double),
                ("aux_azimuth", ctypes.c_double), ("aux_tilt", ctypes.c_double), ("aux_roll", ctypes.c_double),
                ("aux_zoom", ctypes.c_double), ("main_disparity_radius", ctypes.c_double),
                ("aux_disparity_radius", ctypes.c_double)]


class TexParams(ctypes.Structure):
    _fields_ = [("min_shot", ctypes.c_float), ("scale_shot", ctypes.c_float),
                ("diff_sigma", ctypes.c_float), ("dust_remove", ctypes.c_int32),
                ("corr_magic_scale", ctypes.c_float), ("keep_weights", ctypes.c_int32),
                ("port_offsets", (ctypes.c_float * 2) * 16)]


class PortGeometry(ctypes.Structure):
    """clt_port_geometry: the GeometryCorrection parameters of a quad camera (GC:36-83)."""
    _fields_ = [("num_sensors", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("focal_length", ctypes.c_double), ("pixel_size", ctypes.c_double),
                ("distortion_radius", ctypes.c_double), ("distortion", ctypes.c_double * 7),
                ("rxy", (ctypes.c_double * 2) * 4), ("pxy0", (ctypes.c_double * 2) * 4),
                ("roll_deg", ctypes.c_double * 4), ("extrinsic", ctypes.c_double * 13)]


def default_port_geometry(width=2592, height=1936, **kw):
    """Zero distortion and extrinsics, rXY of GC:78, sensor centres at the frame centre,
    focal length 4.5 mm, pixel 2.2 um, distortion radius 2.8512 mm (GC:37-39)."""
    p = PortGeometry()
    p.num_sensors, p.width, p.height = 4, width, height
    p.focal_length, p.pixel_size, p.distortion_radius = 4.5, 2.2, 2.8512
    for i, (x, y) in enumerate(QUAD_PORT_OFFSETS):
        p.rxy[i][0], p.rxy[i][1] = x, y
        p.pxy0[i][0], p.pxy0[i][1] = 0.5 * width, 0.5 * height
    for k, v in kw.items():
        if k in ("distortion", "extrinsic", "roll_deg"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        elif k in ("rxy", "pxy0"):
            arr = getattr(p, k)
            for i, (x, y) in enumerate(v):
                arr[i][0], arr[i][1] = x, y
        else:
            setattr(p, k, v)
    return p


# Reference quad pairs, Correlation2d.java:52-58 {first, second, orientation, scale}
QUAD_PAIRS = [(0, 1, 0, 1), (2, 3, 0, 1), (0, 2, 1, 1), (1, 3, 1, 1), (0, 3, 2, 1), (2, 1, 3, 1)]
# ImageDtt.java:1648-1652 port_offsets of a quad camera
QUAD_PORT_OFFSETS = [(-0.5, -0.5), (0.5, -0.5), (-0.5, 0.5), (0.5, 0.5)]


def default_corr_params(**kw):
    """ImageDttParameters / CLTParameters defaults (IDP:44-73, ECP:2529-2558)."""
    p = CorrParams(fat_zero=0.0, corr_red=0.5, corr_blue=0.2, corr_offset=0.1, twice_diagonal=1,
                   min_corr=0.02, strip_height=9, enhortho_width=2.0, enhortho_scale=0.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def default_corr_lma_params(**kw):
    """ImageDttParameters defaults of the per-tile LMA and foreground correction (IDP:32-110)."""
    p = CorrLmaParams(enable=1, mix_corr_poly=1, min_poly_strength=0.2, ortho_vasw_pwr=2.0,
                      cnvx_hwnd_size=4, cnvx_add3x3=1, cnvx_weight=0.5,
                      lma_adjust_wm=1, lma_adjust_wy=0, lma_adjust_wxy=1, lma_adjust_ag=1,
                      lma_half_width=2.0, lma_cost_wy=0.1, lma_cost_wxy=0.1, lma_lambda_initial=0.1,
                      lma_lambda_scale_good=0.5, lma_lambda_scale_bad=8.0, lma_lambda_max=100.0,
                      lma_rms_diff=0.001, lma_num_iter=20, fo_correct=1, fo_far=0, fo_ortho=1,
                      fo_min_strength=0.1, fo_min_eff=0.25, fo_min_eff_ratio=0.65, fo_max_hwidth=3.0,
                      fo_min_diff=0.2, fo_overcorrection=0.0, fo_lim_overcorr=2.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


DISP_LAYERS = 12
BI_LAYERS = 26
(BI_DISP_FULL, BI_DISP_HOR, BI_DISP_VERT, BI_DISP_DIAGM, BI_DISP_DIAGO,
 BI_ADISP_FULL, BI_ADISP_HOR, BI_ADISP_VERT, BI_ADISP_DIAGM, BI_ADISP_DIAGO,
 BI_DISP_CROSS, BI_DISP_CROSS_DX, BI_DISP_CROSS_DY,
 BI_STR_FULL, BI_STR_HOR, BI_STR_VERT, BI_STR_DIAGM, BI_STR_DIAGO,
 BI_ASTR_FULL, BI_ASTR_HOR, BI_ASTR_VERT, BI_ASTR_DIAGM, BI_ASTR_DIAGO,
 BI_STR_CROSS, BI_STR_ALL, BI_TARGET) = range(BI_LAYERS)


def default_rig_params(**kw):
    """BiQuadParameters and ImageDttParameters defaults of the rig path (BQP:32-83, IDP:81-112)."""
    p = RigParams(use_poly=1, use_xy_poly=1, no_int_x0=1, notch_mode=0, lt_rad=0, corr_strip_notch=11,
                  corr_notch_hwidth=2.0, corr_notch_blur=2.0, cm_max_normalization=0.55, max_corr_radius=3.9)
    for k, v in kw.items():
        setattr(p, k, v)
    return p
(DISP_INT, DISP_CM, DISP_STRENGTH, DISP_HOR, DISP_HOR_STR, DISP_VERT, DISP_VERT_STR, DISP_CM_WEIGHT,
 DISP_POLY, DISP_POLY_STR, DISP_CM_RAW, DISP_STR_RAW) = range(DISP_LAYERS)


def default_tex_params(num_cams=4, port_offsets=None, **kw):
    """CLTParameters defaults (ECP:2584-2594) and the quad port offsets (ID:1648-1652); other camera
    counts default to a circle of radius 0.5 as clt_default_tex_params does (the reference has only
    the quad's)."""
    p = TexParams(min_shot=10.0, scale_shot=3.0, diff_sigma=5.0, dust_remove=1,
                  corr_magic_scale=0.85, keep_weights=1)
    if port_offsets is not None:
        offs = port_offsets
    elif num_cams == 4:
        offs = QUAD_PORT_OFFSETS
    else:
        offs = [(0.5 * math.cos(2 * math.pi * i / num_cams), 0.5 * math.sin(2 * math.pi * i / num_cams))
                for i in range(num_cams)]
    for i, (x, y) in enumerate(offs):
        p.port_offsets[i][0] = x
        p.port_offsets[i][1] = y
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def build():
    """Compile the oracle libraries with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_LIBS = {}
_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int32)
_VP = ctypes.c_void_p


def lib(precision=64):
    """Load libclt_oracle64.so (FP64 restatement) or libclt_oracle32.so (FP32 port)."""
    if precision in _LIBS:
        return _LIBS[precision]
    path = os.path.join(_HERE, "libclt_oracle%d.so" % precision)
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    L.orc_real_size.restype = ctypes.c_int
    L.orc_hwindow.argtypes = [_D]
    L.orc_tables.argtypes = [_VP, _VP, _VP]
    L.orc_dttt.argtypes = [ctypes.c_int, _D, ctypes.c_int, _D]
    L.orc_fold_tile.argtypes = [_D, ctypes.c_double, ctypes.c_double, ctypes.c_int, _D]
    L.orc_unfold_tile.argtypes = [_D, ctypes.c_int, _D]
    L.orc_lpf.argtypes = [ctypes.c_double, _D]
    L.orc_fract_shift.argtypes = [_D, ctypes.c_double, ctypes.c_double]
    L.orc_convolve_tile.argtypes = [_D, _D]
    L.orc_convert_one.argtypes = [ctypes.POINTER(Geom), _F, ctypes.c_int, _VP, _VP, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_int, _D, _D]
    L.orc_iclt_tile.argtypes = [_D, _VP, _D]
    L.orc_corr_unfold.argtypes = [_D, _D]
    L.orc_strip.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]
    L.orc_combine_strips.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _D]
    L.orc_max_xy_int_axis.argtypes = [_D, ctypes.c_double, _I]
    L.orc_half_flat_top_window.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_double, _D]
    L.orc_max_x_cm.argtypes = [_D, ctypes.c_int, ctypes.c_int, _D]
    L.orc_max_xs_ortho.argtypes = [_D, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double, _D]
    L.orc_tile_debayer.argtypes = [ctypes.c_int, _D, _D]
    L.orc_tile_combine_rgba.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, ctypes.POINTER(TexParams),
                                        ctypes.c_int, _D, _VP]
    L.orc_dust_remove.argtypes = [_D, ctypes.c_int]
    L.orc_convert_frame.argtypes = [ctypes.POINTER(Geom), _VP, ctypes.c_int, _VP, _VP, _VP,
                                    ctypes.c_int, ctypes.c_int, _VP, ctypes.c_int, _D, _VP, ctypes.c_int]
    L.orc_imclt_frame.argtypes = [ctypes.POINTER(Geom), _D, _VP, _VP, _D, ctypes.c_int]
    L.orc_corr_frame.argtypes = [ctypes.POINTER(Geom), _D, _VP, ctypes.c_int, ctypes.POINTER(Pair),
                                 ctypes.c_int, ctypes.POINTER(CorrParams), ctypes.POINTER(CorrLmaParams),
                                 ctypes.c_double, _D, _VP, _VP, ctypes.c_int]
    L.orc_corr_lma_frame.argtypes = [ctypes.POINTER(Pair), ctypes.c_int, ctypes.POINTER(CorrLmaParams), _F,
                                     ctypes.c_int, _D, ctypes.c_int]
    L.orc_corr_lma.argtypes = [_D, ctypes.POINTER(Pair), ctypes.c_int, ctypes.c_char_p, ctypes.c_double,
                               ctypes.POINTER(CorrLmaParams), _D]
    L.orc_filter_convex.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, _D]
    L.orc_foreground_correct.argtypes = [ctypes.POINTER(CorrLmaParams), _I, _D, ctypes.c_double]
    L.orc_prepare_kernels.argtypes = [_F, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      _VP, ctypes.c_int, _F, _F]
    L.orc_foreground_correct.restype = ctypes.c_double
    L.orc_textures_frame.argtypes = [ctypes.POINTER(Geom), _D, _VP, ctypes.c_int,
                                     ctypes.POINTER(TexParams), ctypes.POINTER(CorrParams),
                                     ctypes.c_double, _VP, _D, _VP, ctypes.c_int]
    L.orc_combine_textures.argtypes = [ctypes.POINTER(Geom), _VP, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, _D]
    L.orc_rot_matrices.argtypes = [ctypes.POINTER(PortGeometry), _D, _VP]
    L.orc_r_by_rdist.argtypes = [ctypes.POINTER(PortGeometry), ctypes.c_double]
    L.orc_r_by_rdist.restype = ctypes.c_double
    L.orc_port_coords.argtypes = [ctypes.POINTER(PortGeometry), ctypes.c_int, _D, _D, _D, _D, _VP]
    L.orc_rig_analysis.argtypes = [ctypes.POINTER(Geom), _VP, ctypes.c_int, ctypes.POINTER(CorrParams),
                                   ctypes.POINTER(CorrLmaParams), ctypes.POINTER(RigParams), _D, _D, _D, _D, _D,
                                   _VP, _D, ctypes.c_int]
    L.orc_rig_frame.argtypes = [ctypes.POINTER(Geom), _D, _VP, ctypes.c_int, ctypes.POINTER(CorrParams),
                                ctypes.POINTER(CorrLmaParams), ctypes.POINTER(RigParams), ctypes.c_double, _VP,
                                _D, _D, _D, _D, _D, _D, ctypes.c_int]
    L.orc_port_coords_rig.argtypes = [ctypes.POINTER(PortGeometry), ctypes.POINTER(PortGeometry),
                                      ctypes.POINTER(RigOffset), ctypes.c_int, _D, _D, _D, _D]
    L.orc_quadratic_max.argtypes = [_D, _D, _D, _D, ctypes.c_int, _D]
    L.orc_single2d_cm.argtypes = [_D, ctypes.c_double, ctypes.c_double, _D]
    L.orc_corr_peak_offsets.argtypes = [_D, ctypes.c_int, ctypes.c_double, _D]
    L.orc_single2d_poly.argtypes = [_D, ctypes.c_double, ctypes.POINTER(CorrLmaParams), _D]
    _LIBS[precision] = L
    return L


def _dp(a):
    return a.ctypes.data_as(_D)


def _vp(a):
    return None if a is None else a.ctypes.data_as(_VP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def make_geom(width, height, num_cams=4, num_colors=3, red_row=0, red_col=1, kernels=True,
              kernel_step=16):
    kh, kv = (width // kernel_step + 2, height // kernel_step + 2) if kernels else (0, 0)
    return Geom(width=width, height=height, num_cams=num_cams, num_colors=num_colors,
                red_row=red_row, red_col=red_col, kernels_hor=kh, kernels_vert=kv,
                kernel_step=kernel_step, max_tasks=0)


# ---------------------------------------------------------------- primitives
def hwindow(precision=64):
    out = np.zeros(8)
    lib(precision).orc_hwindow(_dp(out))
    return out


def tables(precision=64):
    fi = np.zeros((64, 4), np.int32)
    ui = np.zeros(256, np.int32)
    uk = np.zeros((4, 256))
    lib(precision).orc_tables(fi.ctypes.data_as(_VP), ui.ctypes.data_as(_VP), uk.ctypes.data_as(_VP))
    return fi, ui, uk


def dttt(which, x, mode, precision=64):
    """which: 'iv' (DR:503 dttt_iv), 'iie' (DR:596), 'iiie' (DR:648); x: 64 values."""
    x = _f64(x).reshape(64)
    y = np.zeros(64)
    lib(precision).orc_dttt({"iv": 0, "iie": 1, "iiie": 2}[which], _dp(x), mode, _dp(y))
    return y


def fold_tile(x256, shift_h, shift_v, mode, precision=64):
    x = _f64(x256).reshape(256)
    y = np.zeros(64)
    lib(precision).orc_fold_tile(_dp(x), shift_h, shift_v, mode, _dp(y))
    return y


def unfold_tile(x64, mode, precision=64):
    x = _f64(x64).reshape(64)
    y = np.zeros(256)
    lib(precision).orc_unfold_tile(_dp(x), mode, _dp(y))
    return y


def lpf(sigma, precision=64):
    out = np.zeros(64)
    lib(precision).orc_lpf(sigma, _dp(out))
    return out


def fract_shift(clt4x64, sx, sy, precision=64):
    c = _f64(clt4x64).reshape(256).copy()
    lib(precision).orc_fract_shift(_dp(c), sx, sy)
    return c.reshape(4, 64)


def convolve_tile(data4x64, kern4x64, precision=64):
    d = _f64(data4x64).reshape(256).copy()
    k = _f64(kern4x64).reshape(256)
    lib(precision).orc_convolve_tile(_dp(d), _dp(k))
    return d.reshape(4, 64)


def convert_one(geom, image, color, cx, cy, kernels=None, offsets=None, do_shift=True, precision=64):
    """One camera/colour tile: returns (clt[4][64] in the CPU [h][v] order, residual[2])."""
    image = np.ascontiguousarray(image, np.float32)
    kern = None if kernels is None else np.ascontiguousarray(kernels, np.float32)
    offs = None if offsets is None else np.ascontiguousarray(offsets, np.float32)
    clt = np.zeros(256)
    res = np.zeros(2)
    lib(precision).orc_convert_one(ctypes.byref(geom), image.ctypes.data_as(_F), image.shape[-1],
                                   _vp(kern), _vp(offs), color, cx, cy, int(do_shift), _dp(clt), _dp(res))
    return clt.reshape(4, 64), res


def iclt_tile(clt4x64, lpf64=None, precision=64):
    c = _f64(clt4x64).reshape(256)
    l = None if lpf64 is None else _f64(lpf64)
    out = np.zeros(256)
    lib(precision).orc_iclt_tile(_dp(c), _vp(l), _dp(out))
    return out.reshape(16, 16)


def corr_unfold(q4x64, precision=64):
    q = _f64(q4x64).reshape(256)
    out = np.zeros(225)
    lib(precision).orc_corr_unfold(_dp(q), _dp(out))
    return out.reshape(15, 15)


def strip(corr225, hwidth, direction, ss=1, precision=64):
    c = _f64(corr225).reshape(225)
    out = np.zeros(hwidth * 15)
    lib(precision).orc_strip(_dp(c), hwidth, direction, ss, _dp(out))
    return out.reshape(hwidth, 15)


def combine_strips(strips, offset=0.1, twice_diagonal=True, precision=64):
    s = _f64(strips)
    n, length = s.shape[0], s[0].size
    out = np.zeros(length)
    lib(precision).orc_combine_strips(_dp(s), n, length, offset, int(twice_diagonal), _dp(out))
    return out.reshape(s.shape[1:])


def max_xy_int_axis(data15, min_max=0.02, precision=64):
    d = _f64(data15).reshape(-1)[:15].copy()
    ix = ctypes.c_int32(0)
    ok = lib(precision).orc_max_xy_int_axis(_dp(d), min_max, ctypes.byref(ix))
    return ix.value if ok else None


def half_flat_top_window(ihwidth, hwidth, blur, normalize=True, notch=False, scale=1.0):
    out = np.zeros(ihwidth)
    lib().orc_half_flat_top_window(ihwidth, hwidth, blur, int(normalize), int(notch), scale, _dp(out))
    return out


def max_x_cm(strip_data, ixcenter, precision=64):
    s = _f64(strip_data)
    out = np.zeros(3)
    ok = lib(precision).orc_max_x_cm(_dp(s.reshape(-1)), s.shape[0], ixcenter, _dp(out))
    return out if ok else None


def max_xs_ortho(corrs, offset=0.1, is_vert=False, enhortho_width=2.0, enhortho_scale=0.0, precision=64):
    c = _f64(corrs).reshape(-1, 225)
    out = np.zeros(2)
    ok = lib(precision).orc_max_xs_ortho(_dp(c), c.shape[0], offset, int(is_vert), enhortho_width,
                                         enhortho_scale, _dp(out))
    return out if ok else None


def tile_debayer(rb, tile, precision=64):
    t = _f64(tile).reshape(256)
    out = np.zeros(256)
    lib(precision).orc_tile_debayer(int(rb), _dp(t), _dp(out))
    return out.reshape(16, 16)


def tile_combine_rgba(deb, iclt, lt_window, port_mask=None, tex_params=None, chn_weights=None, precision=64):
    """tile_combine_rgba (ID:2830-3000) on the debayered tiles deb [ports][3][256] followed by the
    final mix (ID:2789-2801) of iclt [ports][3][256]; lt_window [256]; colour weights R, B, G
    (default: corr_red 0.5, corr_blue 0.2, ID:1571-1581).  Returns (tile [planes][256], stats
    [4*ports]) with planes = 4 + ports + 4 when tex_params.keep_weights, else 4."""
    deb = _f64(deb).reshape(-1, 3, 256)
    ports = deb.shape[0]
    tp = tex_params or default_tex_params(ports)
    mask = (1 << ports) - 1 if port_mask is None else port_mask
    cw = _f64(np.array([0.5, 0.2, 1.0]) / 1.7 if chn_weights is None else chn_weights)
    planes = 4 + ports + 4 if tp.keep_weights else 4
    out = np.zeros((planes, 256))
    stats = np.zeros(4 * ports)
    lib(precision).orc_tile_combine_rgba(ports, mask, _dp(deb), _dp(_f64(iclt).reshape(ports, 3, 256)),
                                         _dp(_f64(lt_window).reshape(256)), _dp(cw), ctypes.byref(tp), planes,
                                         _dp(out), _dp(stats))
    return out, stats


def dust_remove(pw):
    """dust_remove (ID:3002-3019) of port weights [ports][256]; returns the new weights."""
    w = np.array(pw, np.float64, copy=True, order="C").reshape(len(pw), -1)
    assert w.shape[1] == 256
    lib().orc_dust_remove(_dp(w), w.shape[0])
    return w


# ---------------------------------------------------------------- frame level
def _ptr_array(arrs):
    if arrs is None:
        return None, None
    keep = [np.ascontiguousarray(a, np.float32) for a in arrs]
    ptrs = (ctypes.c_void_p * len(keep))(*[k.ctypes.data for k in keep])
    return ptrs, keep


def convert_frame(geom, images, tasks, kernels=None, offsets=None, lpf_mask=0, lpf3=None,
                  no_fract_shift=False, nthreads=None, precision=64, want_residual=False):
    """ID clt_aberrations_quad_corr convert stage over a task list.

    images: [cams][H][W] float32 mosaics; kernels/offsets: per camera arrays in the GPU upload
    layout or None; tasks: int32 [ntasks][2+2C] records.  Returns CLT float64
    [cams][tilesY*tilesX][colors][4][8][8] ([v][h] blocks) and, optionally, residual shifts."""
    nthreads = nthreads or os.cpu_count()
    C, K = geom.num_cams, geom.num_colors
    nt = (geom.width // 8) * (geom.height // 8)
    img_p, keep_i = _ptr_array(images)
    ker_p, keep_k = _ptr_array(kernels)
    off_p, keep_o = _ptr_array(offsets)
    tasks = np.ascontiguousarray(tasks, np.int32)
    ntasks = tasks.shape[0]
    clt = np.zeros((C, nt, K, 4, 8, 8))
    res = np.zeros((ntasks, C, K, 2)) if want_residual else None
    l3 = None if lpf3 is None else _f64(lpf3).reshape(-1)
    lib(precision).orc_convert_frame(ctypes.byref(geom), ctypes.cast(img_p, _VP),
                                     np.asarray(images).shape[-1],
                                     None if ker_p is None else ctypes.cast(ker_p, _VP),
                                     None if off_p is None else ctypes.cast(off_p, _VP),
                                     tasks.ctypes.data_as(_VP), ntasks, lpf_mask, _vp(l3),
                                     int(no_fract_shift), _dp(clt), _vp(res), nthreads)
    del keep_i, keep_k, keep_o
    return (clt, res) if want_residual else clt


def imclt_frame(geom, clt, valid, lpf64=None, nthreads=None, precision=64):
    """iclt_2d (ID:3559-3666) of every camera/colour: [cams][colors][H+8][W+8] (uncropped)."""
    nthreads = nthreads or os.cpu_count()
    clt = _f64(clt)
    valid = np.ascontiguousarray(valid, np.uint8)
    out = np.zeros((geom.num_cams, geom.num_colors, geom.height + 8, geom.width + 8))
    l = None if lpf64 is None else _f64(lpf64)
    lib(precision).orc_imclt_frame(ctypes.byref(geom), _dp(clt), valid.ctypes.data_as(_VP), _vp(l),
                                   _dp(out), nthreads)
    return out


def corr_frame(geom, clt, tasks, pairs=None, params=None, corr_sigma=0.8, nthreads=None, precision=64,
               lma=None):
    """Pair correlations [task][pair][15][15], combined strip [task][h][15], disparity
    [task][DISP_LAYERS].  lma: CorrLmaParams (default: the reference defaults, enabled)."""
    nthreads = nthreads or os.cpu_count()
    pairs = pairs if pairs is not None else QUAD_PAIRS
    params = params or default_corr_params()
    lma = lma if lma is not None else default_corr_lma_params()
    parr = (Pair * len(pairs))(*[Pair(*p) for p in pairs])
    tasks = np.ascontiguousarray(tasks, np.int32)
    ntasks = tasks.shape[0]
    clt = _f64(clt)
    corr = np.zeros((ntasks, len(pairs), 15, 15))
    combo = np.zeros((ntasks, params.strip_height, 15))
    disp = np.zeros((ntasks, DISP_LAYERS))
    rc = lib(precision).orc_corr_frame(ctypes.byref(geom), _dp(clt), tasks.ctypes.data_as(_VP), ntasks,
                                       parr, len(pairs), ctypes.byref(params), ctypes.byref(lma), corr_sigma,
                                       _dp(corr), _vp(combo), _vp(disp), nthreads)
    if rc:
        raise ValueError("orc_corr_frame failed: %d" % rc)
    return corr, combo, disp


def corr_lma_frame(corrs, disp, pairs=None, lma=None, nthreads=None):
    """The per-tile LMA chain (ID:2254-2336) alone on given pair correlations [T][P][15][15]
    (float32) and disparity records [T][DISP_LAYERS]; returns the updated records (float64)."""
    nthreads = nthreads or os.cpu_count()
    pairs = pairs if pairs is not None else QUAD_PAIRS
    lma = lma if lma is not None else default_corr_lma_params()
    parr = (Pair * len(pairs))(*[Pair(*p) for p in pairs])
    corrs = np.ascontiguousarray(corrs, np.float32)
    out = np.array(disp, np.float64, copy=True, order="C")
    assert corrs.shape[0] == out.shape[0] and out.shape[1] == DISP_LAYERS
    rc = lib(64).orc_corr_lma_frame(parr, len(pairs), ctypes.byref(lma), corrs.ctypes.data_as(_F),
                                    out.shape[0], _dp(out), nthreads)
    if rc:
        raise ValueError("orc_corr_lma_frame failed: %d" % rc)
    return out


def corr_lma(corrs, xcenter, sel=None, pairs=None, lma=None):
    """One Correlations2dLMA fit (C2:1754-1900) over the selected pairs of one tile
    (corrs [P][15][15]); returns (disparity, strength, Wm, Wxy) or None when the fit fails."""
    pairs = pairs if pairs is not None else QUAD_PAIRS
    lma = lma if lma is not None else default_corr_lma_params()
    sel = [1] * len(pairs) if sel is None else sel
    parr = (Pair * len(pairs))(*[Pair(*p) for p in pairs])
    c = _f64(np.asarray(corrs, np.float64).reshape(len(pairs), 225))
    out = np.zeros(4)
    ok = lib(64).orc_corr_lma(_dp(c), parr, len(pairs), bytes(bytearray(int(bool(x)) for x in sel)),
                              float(xcenter), ctypes.byref(lma), _dp(out))
    return tuple(out) if ok == 1 else None


def filter_convex(data15, hwin, x0c, y0c, add3x3=True, nc_cost=0.5):
    """filterConvex (C2:2029-2207): [15][15] weights of the cells selected around (x0c, y0c)."""
    d = _f64(np.asarray(data15, np.float64).reshape(225))
    w = np.zeros(225)
    lib(64).orc_filter_convex(_dp(d), hwin, x0c, y0c, int(add3x3), nc_cost, _dp(w))
    return w.reshape(15, 15)


def foreground_correct(dirs, full_disp, lma=None):
    """foregroundCorrect (C2:1663-1750); dirs: 4 entries (disparity, strength, Wm, Wxy) or None."""
    lma = lma if lma is not None else default_corr_lma_params()
    have = np.array([d is not None for d in dirs], np.int32)
    dds = np.array([d if d is not None else (0, 0, 0, 0) for d in dirs], np.float64)
    return lib(64).orc_foreground_correct(ctypes.byref(lma), have.ctypes.data_as(_I), _dp(dds), full_disp)


def textures_frame(geom, clt, tasks, tex_params=None, corr_params=None, corr_sigma=0.8,
                   extra_disparity=None, nthreads=None, precision=64):
    """Texture tiles [task][planes][16][16] and stats [task][4*cams]."""
    nthreads = nthreads or os.cpu_count()
    tp = tex_params or default_tex_params(geom.num_cams)
    cp = corr_params or default_corr_params()
    tasks = np.ascontiguousarray(tasks, np.int32)
    ntasks = tasks.shape[0]
    planes = 4 + geom.num_cams + 4 if tp.keep_weights else 4
    tiles = np.zeros((ntasks, planes, 16, 16))
    stats = np.zeros((ntasks, 4 * geom.num_cams))
    ed = None if extra_disparity is None else _f64(extra_disparity)
    rc = lib(precision).orc_textures_frame(ctypes.byref(geom), _dp(_f64(clt)), tasks.ctypes.data_as(_VP),
                                           ntasks, ctypes.byref(tp), ctypes.byref(cp), corr_sigma, _vp(ed),
                                           _dp(tiles), _vp(stats), nthreads)
    if rc:
        raise ValueError("orc_textures_frame failed: %d" % rc)
    return tiles, stats


def combine_textures(geom, tasks, tiles, overlap=True, sharp_alpha=False):
    """combineRGBATiles (ID:3855-3991): [planes][H][W] with overlap, else the [planes][2H][2W]
    mosaic of whole tiles; sharp_alpha: alpha from the centre 8x8 of each tile only."""
    tasks = np.ascontiguousarray(tasks, np.int32)
    tiles = _f64(tiles)
    planes = tiles.shape[1]
    k = 1 if overlap else 2
    out = np.zeros((planes, k * geom.height, k * geom.width))
    lib().orc_combine_textures(ctypes.byref(geom), tasks.ctypes.data_as(_VP), tasks.shape[0], _dp(tiles),
                               planes, int(bool(overlap)), int(bool(sharp_alpha)), _dp(out))
    return out


# ------------------------------------------------------------------ port geometry (always FP64)
def rot_matrices(pg):
    """Per sensor rotation/zoom matrices [4][3][3] and their derivatives by azimuth, tilt, roll,
    zoom [4][4][3][3] (GC:1043-1147)."""
    r = np.zeros((4, 3, 3))
    d = np.zeros((4, 4, 3, 3))
    lib(64).orc_rot_matrices(ctypes.byref(pg), _dp(r), _vp(d))
    return r, d


def r_by_rdist(pg, rdist):
    """Non-distorted / distorted radius ratio at relative distorted radius rdist (GC:2802-2830)."""
    return lib(64).orc_r_by_rdist(ctypes.byref(pg), float(rdist))


def port_coords(pg, px, py, disparity, deriv=False):
    """Port centres [n][4][2] of tile centres (px, py) at disparity (GC:2159-2314); with deriv also
    [n][8][14]: d/d 13 extrinsics (pXYderiv) and d/d disparity."""
    px = np.ascontiguousarray(np.atleast_1d(px), np.float64)
    py = np.ascontiguousarray(np.broadcast_to(py, px.shape), np.float64)
    dd = np.ascontiguousarray(np.broadcast_to(disparity, px.shape), np.float64)
    xy = np.zeros((len(px), 4, 2))
    dv = np.zeros((len(px), 8, 14)) if deriv else None
    lib(64).orc_port_coords(ctypes.byref(pg), len(px), _dp(px), _dp(py), _dp(dd), _dp(xy),
                            _vp(dv) if deriv else None)
    return (xy, dv) if deriv else xy


def rig_frame(geom, clt, tasks, target=None, params=None, lma=None, rig=None, corr_sigma=0.8, nthreads=None):
    """clt_bi_quad (ID:8731-9504) correlation part over a converted 8-camera frame (cameras 0..3 the
    main quad, 4..7 the auxiliary one).  Returns a dict: corr_main / corr_aux [T][6][15][15],
    disp_main / disp_aux [T][DISP_LAYERS] (correlate2D records, no LMA), inter [T][15][15],
    bimap [T][BI_LAYERS]."""
    nthreads = nthreads or os.cpu_count()
    params = params or default_corr_params()
    lma = lma if lma is not None else default_corr_lma_params()
    rig = rig if rig is not None else default_rig_params()
    tasks = np.ascontiguousarray(tasks, np.int32)
    T = tasks.shape[0]
    out = {"corr_main": np.zeros((T, 6, 15, 15)), "disp_main": np.zeros((T, DISP_LAYERS)),
           "corr_aux": np.zeros((T, 6, 15, 15)), "disp_aux": np.zeros((T, DISP_LAYERS)),
           "inter": np.zeros((T, 15, 15)), "bimap": np.zeros((T, BI_LAYERS))}
    tg = None if target is None else np.ascontiguousarray(target, np.float32)
    rc = lib(64).orc_rig_frame(ctypes.byref(geom), _dp(_f64(clt)), tasks.ctypes.data_as(_VP), T, ctypes.byref(params),
                               ctypes.byref(lma), ctypes.byref(rig), corr_sigma, _vp(tg),
                               *[_dp(out[k]) for k in ("corr_main", "disp_main", "corr_aux", "disp_aux", "inter",
                                                       "bimap")], nthreads)
    if rc:
        raise ValueError("orc_rig_frame failed: %d" % rc)
    return out


def rig_analysis(geom, tasks, corr_main, disp_main, corr_aux, disp_aux, inter, target=None, params=None, lma=None,
                 rig=None, nthreads=None):
    """The rig analysis alone (tileCorrs of both quads, tileInterCamCorrs, the low-texture pass) on
    given correlations and correlate2D records; returns bimap [T][BI_LAYERS]."""
    nthreads = nthreads or os.cpu_count()
    params = params or default_corr_params()
    lma = lma if lma is not None else default_corr_lma_params()
    rig = rig if rig is not None else default_rig_params()
    tasks = np.ascontiguousarray(tasks, np.int32)
    T = tasks.shape[0]
    arrs = [_f64(a) for a in (corr_main, disp_main, corr_aux, disp_aux, inter)]
    bimap = np.zeros((T, BI_LAYERS))
    tg = None if target is None else np.ascontiguousarray(target, np.float32)
    rc = lib(64).orc_rig_analysis(ctypes.byref(geom), tasks.ctypes.data_as(_VP), T, ctypes.byref(params),
                                  ctypes.byref(lma), ctypes.byref(rig), *[_dp(a) for a in arrs], _vp(tg), _dp(bimap),
                                  nthreads)
    if rc:
        raise ValueError("orc_rig_analysis failed: %d" % rc)
    return bimap


def port_coords_rig(main, aux, rig, px, py, disparity):
    """Port centres of both quads of a rig (ID:8970-8993): [n][8][2], main ports then auxiliary."""
    px, py, d = (_f64(np.atleast_1d(a)) for a in (px, py, disparity))
    xy = np.zeros((len(px), 8, 2))
    lib(64).orc_port_coords_rig(ctypes.byref(main), ctypes.byref(aux), ctypes.byref(rig), len(px), _dp(px), _dp(py),
                                _dp(d), _dp(xy))
    return xy


def quadratic_max(x, y, z, w):
    """PA:201-224 over PA:350-580: (x, y, f, A, B) of the weighted quadratic fit, or None."""
    x, y, z, w = (_f64(a) for a in (x, y, z, w))
    out = np.zeros(5)
    ok = lib(64).orc_quadratic_max(_dp(x), _dp(y), _dp(z), _dp(w), len(x), _dp(out))
    return out if ok else None


def single2d_cm(corr225, xcenter=0.0, radius=3.9):
    out = np.zeros(3)
    lib(64).orc_single2d_cm(_dp(_f64(corr225).ravel()), xcenter, radius, _dp(out))
    return out


def corr_peak_offsets(corr, radius=3.9):
    """[n][6][225] quad-pair correlations -> [n][6][3] (dx, dy, value at the maximum): centre of
    mass of the positive values within radius of each pair's integer maximum (mismatchPairsCM,
    C2:1336-1430, as the clt_mismatch measurement, ID:2342-2420)."""
    c = np.ascontiguousarray(_f64(corr).reshape(-1, 6, 225))
    out = np.zeros((c.shape[0], 6, 3))
    lib(64).orc_corr_peak_offsets(_dp(c), c.shape[0], float(radius), _dp(out))
    return out


def single2d_poly(corr225, xcenter=0.0, lma=None):
    lma = lma if lma is not None else default_corr_lma_params()
    out = np.zeros(3)
    lib(64).orc_single2d_poly(_dp(_f64(corr225).ravel()), xcenter, ctypes.byref(lma), _dp(out))
    return out


def prepare_kernels(raw, kernels_hor, kernels_vert, ks=64, kernel_step=16, sensor_xy=None, norm_kern=True):
    """calculateCLTKernel (QC:433-668) of one camera: raw kernel stack [colors][KY*ks][KX*ks] ->
    (kernels [KY][KX][colors][4][8][8] upload layout [v][h], offsets [KY][KX][colors][8]), float32."""
    raw = np.ascontiguousarray(raw, np.float32)
    colors = raw.shape[0]
    assert raw.shape == (colors, kernels_vert * ks, kernels_hor * ks)
    kern = np.zeros((kernels_vert, kernels_hor, colors, 4, 8, 8), np.float32)
    offs = np.zeros((kernels_vert, kernels_hor, colors, 8), np.float32)
    sxy = None if sensor_xy is None else np.ascontiguousarray(sensor_xy, np.float32)
    rc = lib(64).orc_prepare_kernels(raw.ctypes.data_as(_F), ks, kernels_hor, kernels_vert, colors, kernel_step,
                                     None if sxy is None else sxy.ctypes.data_as(_VP), int(bool(norm_kern)),
                                     kern.ctypes.data_as(_F), offs.ctypes.data_as(_F))
    if rc:
        raise ValueError("orc_prepare_kernels failed: %d" % rc)
    return kern, offs
