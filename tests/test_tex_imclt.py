"""clt_exec_textures_imclt: textures and imclt_rbg in one call must give the outputs of
clt_exec_textures followed by clt_exec_imclt (both checked against the oracle in
test_gpu_parity.py).  The texture tiles and statistics run the same per-tile code on the same
inputs and must be bit-identical; the image planes are the same four overlapping tile halves summed
in another order (two column launches instead of a strip sweep), so they agree to float rounding.
Cases: dense and sparse task lists, even and odd tile columns, segment boundaries, and the inputs
that make the call fall back to the two passes (extra disparity, no LPF, a zero-flag task, two
tasks on one tile)."""
import numpy as np
import pytest

from helpers import TEX_TOL, gpu_clt_cpu_layout, gpu_processor, log, oracle_geom, rel_rms, tex_planes, valid_map

pytestmark = pytest.mark.gpu


def _outputs(tp, sc):
    tiles, stats = tp.getTextures()
    img = np.stack([tp.getRBG(c) for c in range(sc.num_cams)])
    return tiles, stats, img


def _compare(pk, sc, tasks=None, extra=None, apply_lpf=True, keep_weights=1):
    def run(fused):
        tp = gpu_processor(pk, sc)
        if tasks is not None:
            tp.setTasks(tasks)
        tp.setCorrSigma(0.8)
        tp.setTexParams(keep_weights=keep_weights)
        tp.execConverCorrectTiles()
        tp.setTiming(True)
        if fused:
            tp.execTexturesImclt(apply_lpf, extra)
        else:
            tp.execTextures(extra)
            tp.execImcltRbg(apply_lpf)
        out = _outputs(tp, sc) + (tp.getTiming("tex_imclt")[1],)
        tp.close()
        return out

    (t1, s1, i1, _), (t2, s2, i2, one_pass) = run(False), run(True)
    log("  one pass: %s" % bool(one_pass))
    tiles_same = np.array_equal(t1.view(np.int32), t2.view(np.int32))
    stats_same = np.array_equal(s1.view(np.int32), s2.view(np.int32))
    scale = np.abs(i1).max()
    img_err = float(np.abs(i1.astype(np.float64) - i2).max() / scale)
    log("  tiles bit-identical=%s stats bit-identical=%s image max |diff| / max = %.3g" %
        (tiles_same, stats_same, img_err))
    return tiles_same, stats_same, img_err, bool(one_pass), (i1, i2)


def test_dense_quad(pk):
    log("case: dense quad 128x96 (16 x 12 tiles, even column count)")
    sc = pk.synth.make_scene(128, 96, seed=21, disparity=("field", 0.0, 6.0))
    ts, ss, ie, one, _ = _compare(pk, sc)
    assert one and ts and ss and ie < 1e-6


def test_sparse_odd_columns_segments(pk):
    log("case: sparse tasks, 136x152 (17 x 19 tiles: odd columns, several 8-row segments)")
    sc = pk.synth.make_scene(136, 152, seed=22, disparity=("field", 0.0, 5.0), task_fraction=0.6)
    ts, ss, ie, one, (i1, _) = _compare(pk, sc)
    assert one and ts and ss and ie < 1e-6
    assert np.abs(i1).max() > 0


@pytest.mark.parametrize("cams", [2, 3, 5])
def test_camera_counts(pk, cams):
    log("case: %d cameras on a circle, 96x64" % cams)
    sc = pk.synth.make_scene(96, 64, num_cams=cams, seed=50 + cams, disparity=("field", 0.0, 4.0))
    ts, ss, ie, one, _ = _compare(pk, sc)
    assert one and ts and ss and ie < 1e-6


def test_rgba_only(pk):
    log("case: keep_weights off (4 texture planes)")
    sc = pk.synth.make_scene(128, 96, seed=23, disparity=2.0)
    ts, ss, ie, one, _ = _compare(pk, sc, keep_weights=0)
    assert one and ts and ss and ie < 1e-6


@pytest.mark.parametrize("case", ["extra", "no_lpf", "zero_flag", "duplicate"])
def test_fallbacks_match_two_passes(pk, case):
    log("case: fallback to the two passes (%s)" % case)
    sc = pk.synth.make_scene(128, 96, seed=24, disparity=2.0)
    tasks, extra, lpf = sc.tasks.copy(), None, True
    if case == "extra":
        extra = np.random.default_rng(5).uniform(-1.0, 1.0, len(tasks)).astype(np.float32)
    elif case == "no_lpf":
        lpf = False
    elif case == "zero_flag":
        tasks[7, 0] = 0
    else:
        tasks[5] = tasks[3]   # two tasks on one tile (the same record: convert writes the same CLT)
    ts, ss, ie, one, _ = _compare(pk, sc, tasks=tasks, extra=extra, apply_lpf=lpf)
    assert not one and ts and ss and ie == 0.0


def test_async_side_stream_ordering(pk):
    """imclt_rbg runs on a side stream in async mode: a convert issued right after it (no host
    synchronisation in between) must not change the CLT it reads.  Reference: the same calls in
    synchronous mode."""
    log("case: async imclt_rbg on its side stream, then a convert with another LPF mask")
    sc = pk.synth.make_scene(128, 96, seed=25, disparity=("field", 0.0, 6.0))

    def run(asynchronous):
        tp = gpu_processor(pk, sc)
        tp.setCorrSigma(0.8)
        tp.setLpfRbg(0.7, 0.7, 0.7)
        tp.setAsync(asynchronous)
        tp.execConverCorrectTiles(0)
        tp.execImcltRbg(True)
        tp.execConverCorrectTiles(7)   # rewrites the CLT while imclt_rbg may still run
        img = np.stack([tp.getRBG(c) for c in range(sc.num_cams)])
        clt = tp.getCLT(0)
        tp.close()
        return img, clt

    i0, c0 = run(False)
    i1, c1 = run(True)
    same_img = np.array_equal(i0.view(np.int32), i1.view(np.int32))
    same_clt = np.array_equal(c0.view(np.int32), c1.view(np.int32))
    log("  bit-identical: images=%s, second CLT=%s" % (same_img, same_clt))
    assert same_img and same_clt


@pytest.mark.parametrize("frac", [1.0, 0.6])
def test_one_pass_against_oracle(pk, orc, frac):
    """The one-pass k_tex_imclt outputs against the oracle directly: textures (ID:2662-2812) and the
    LPF'd inverse CLT with overlap-add (ID:3559-3666) of the device CLT (the same input)."""
    log("case: one-pass textures + imclt_rbg vs the FP64 oracle, 136x152, %d %% of the tiles" % (100 * frac))
    sc = pk.synth.make_scene(136, 152, seed=26, disparity=("field", 0.0, 5.0), task_fraction=frac)
    tp = gpu_processor(pk, sc)
    tp.setCorrSigma(0.8)
    tp.execConverCorrectTiles()
    tp.setTiming(True)
    tp.execTexturesImclt(True)
    assert tp.getTiming("tex_imclt")[1] == 1   # the one-pass kernel ran (not the two-pass fallback)
    tiles, stats, img = _outputs(tp, sc)
    clt = gpu_clt_cpu_layout(tp, sc).astype(np.float64)
    tp.close()
    g = oracle_geom(orc, sc)
    o_tiles, o_stats = orc.textures_frame(g, clt, sc.tasks)
    o_img = orc.imclt_frame(g, clt, valid_map(sc), lpf64=orc.lpf(0.8))
    err = {name: rel_rms(tiles[:, sl], o_tiles[:, sl]) for name, sl in tex_planes(4)}
    err["stats"] = rel_rms(stats, o_stats)
    err["image"] = rel_rms(img.reshape(o_img.shape), o_img)
    log("  " + ", ".join("%s=%.3g" % kv for kv in err.items()))
    for name, tol in TEX_TOL.items():
        assert err[name] < tol, name
    assert err["stats"] < 1e-4 and err["image"] < 1e-4


@pytest.mark.parametrize("cams", [3, 4, 5])
def test_lds_layouts_bit_identical(pk, cams):
    """The one-pass kernel's two LDS tile layouts (tuning knob tex_lds_layout: rows 20 floats apart,
    or 16 with swizzled 16-byte column groups) move the same values through LDS: every output is
    bit-identical."""
    log("case: LDS layouts 0 and 1 of the one-pass kernel, %d cameras, 136x152, 70 %% of the tiles" % cams)
    sc = pk.synth.make_scene(136, 152, num_cams=cams, seed=27 + cams, disparity=("field", 0.0, 5.0),
                             task_fraction=0.7)

    def run(layout):
        tp = gpu_processor(pk, sc)
        tp.setCorrSigma(0.8)
        tp.setTexParams(keep_weights=1)
        tp.setTune(tex_lds_layout=layout)
        assert tp.getTune()["tex_lds_layout"] == layout
        tp.execConverCorrectTiles()
        tp.setTiming(True)
        tp.execTexturesImclt(True)
        assert tp.getTiming("tex_imclt")[1] == 1
        out = _outputs(tp, sc)
        tp.close()
        return out

    a, b = run(0), run(1)
    same = [np.array_equal(x.view(np.int32), y.view(np.int32)) for x, y in zip(a, b)]
    log("  bit-identical tiles, stats, images: %s" % same)
    assert all(same)
