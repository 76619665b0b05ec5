"""CPU-side checks of the C ABI library and the host logic (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "clt_tile_processor.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(clt_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(pk):
    lib = pk.load()
    declared = _declared_functions()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(pk._abi.SIGNATURES), "ctypes table out of sync with the header"


def test_struct_layouts_match_oracle(pk, orc):
    for a, b in ((pk._abi.clt_geom, orc.Geom), (pk._abi.clt_pair, orc.Pair),
                 (pk._abi.clt_corr_params, orc.CorrParams), (pk._abi.clt_tex_params, orc.TexParams)):
        assert ctypes.sizeof(a) == ctypes.sizeof(b)
        assert [f[0] for f in a._fields_] == [f[0] for f in b._fields_]


def test_create_without_gpu_or_bad_geometry_fails_cleanly(pk):
    lib = pk.load()
    h = ctypes.c_void_p()
    bad = pk._abi.clt_geom(width=100, height=64, num_cams=4, num_colors=3, red_row=0, red_col=1,
                           kernels_hor=0, kernels_vert=0, kernel_step=16, max_tasks=0)
    assert lib.clt_create(ctypes.byref(bad), ctypes.byref(h)) == -1          # width not a multiple of 8
    bad.width, bad.num_colors = 64, 2
    assert lib.clt_create(ctypes.byref(bad), ctypes.byref(h)) == -1          # 2 colours
    if lib.clt_device_count() == 0:
        good = pk._abi.clt_geom(width=64, height=48, num_cams=4, num_colors=3, red_row=0, red_col=1,
                                kernels_hor=6, kernels_vert=5, kernel_step=16, max_tasks=0)
        assert lib.clt_create(ctypes.byref(good), ctypes.byref(h)) == -4      # CLT_E_NODEVICE
        with pytest.raises(pk.CltError):
            pk.GPUTileProcessor(64, 48)


def test_null_handle_calls_are_rejected(pk):
    lib = pk.load()
    assert lib.clt_destroy(None) == -1
    assert lib.clt_exec_convert(None, 0) == -1
    assert lib.clt_sync(None) == -1


def test_defaults_match_reference_parameters(pk):
    lib = pk.load()
    cp = pk._abi.clt_corr_params()
    lib.clt_default_corr_params(ctypes.byref(cp))
    assert (round(cp.corr_red, 6), round(cp.corr_blue, 6), round(cp.corr_offset, 6)) == (0.5, 0.2, 0.1)
    assert cp.strip_height == 9 and cp.twice_diagonal == 1 and abs(cp.min_corr - 0.02) < 1e-7
    pairs = (pk._abi.clt_pair * 6)()
    assert lib.clt_default_quad_pairs(pairs) == 6
    assert [(p.cam1, p.cam2, p.dir) for p in pairs] == [(0, 1, 0), (2, 3, 0), (0, 2, 1), (1, 3, 1), (0, 3, 2), (2, 1, 3)]
    tp = pk._abi.clt_tex_params()
    lib.clt_default_tex_params(ctypes.byref(tp), 4)
    assert [tuple(tp.port_offsets[i]) for i in range(4)] == [(-0.5, -0.5), (0.5, -0.5), (-0.5, 0.5), (0.5, 0.5)]


def test_task_records_round_trip(pk):
    g = pk.geometry
    c = g.port_centers(64, 48, [1, 7], [0, 5], [2.0, 0.5], g.QUAD_RXY)
    rec = g.make_tasks([1, 7], [0, 5], c, flags=0xFF)
    assert rec.shape == (2, 10) and rec.dtype == np.int32
    assert (rec[1, 1] & 0xFFFF) == 7 and (rec[1, 1] >> 16) == 5
    xy = rec[:, 2:].view(np.float32).reshape(2, 4, 2)
    # zero-distortion closed form: centre - d * rXY (GeometryCorrection.java:2181-2231)
    assert np.allclose(xy[0, :, 0], 12 - 2.0 * g.QUAD_RXY[:, 0])
    assert np.allclose(xy[1, :, 1], 44 - 0.5 * g.QUAD_RXY[:, 1])
    t = pk.TpTask(7, 5, 0.5, 0xFF, xy[1])
    assert np.array_equal(t.as_record(), rec[1])


def test_identity_rotation_matches_closed_form(pk):
    g = pk.geometry
    a = g.port_centers(256, 192, [3, 20], [4, 9], 1.5, g.QUAD_RXY)
    b = g.port_centers(256, 192, [3, 20], [4, 9], 1.5, g.QUAD_RXY, rots=np.tile(np.eye(3), (4, 1, 1)), focal_pix=3000.0)
    assert np.abs(a - b).max() < 1e-9


def test_synthetic_kernels_delta_chain(pk):
    """A near-delta PSF through the kernel chain gives the identity CC kernel (x16 normalisation of
    clt_normalize_kernel, ID:4674-4696)."""
    rng = np.random.default_rng(0)
    k, o = pk.synth.clt_kernels(1, 1, 3, 3, rng, sigma=(0.05, 0.05), jitter=0.0)
    cc = k[0, 1, 1, 0, 0]
    assert abs(cc.mean() - 1.0) < 0.05 and np.abs(k[0, 1, 1, 0, 1:]).max() < 0.05
    assert o.shape == (1, 3, 3, 1, 8)


def test_bayer_phase(pk):
    c = pk.synth.bayer_colors(4, 4, 0, 1)
    assert c[0, 1] == 0 and c[1, 0] == 1 and c[0, 0] == 2 and c[1, 1] == 2   # GR/BG (ref.)
    c = pk.synth.bayer_colors(4, 4, 0, 0)
    assert c[0, 0] == 0 and c[1, 1] == 1 and c[0, 1] == 2                    # RGGB
