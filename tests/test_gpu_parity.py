"""GPU parity of the four kernels against the CPU oracle (FP64 restatement of the reference).

Tolerances (from the acceptance criteria of this path): CLT coefficients and reconstructed images
within 1e-4 relative RMS, correlation-peak disparities within 0.01 px.  The FP32 error of the
oracle's own single-precision build is printed next to the GPU error for scale.
"""
import numpy as np
import pytest

from helpers import (TEX_TOL, gpu_clt_cpu_layout, gpu_processor, log, oracle_geom, rel_rms, task_tiles, tex_planes,
                     valid_map)

pytestmark = pytest.mark.gpu

CLT_TOL = 1e-4
IMG_TOL = 1e-4
DISP_TOL = 0.01


@pytest.fixture(scope="module")
def quad_scene(pk):
    return pk.synth.make_scene(128, 96, seed=11, disparity=("field", 0.0, 6.0))


def _convert_case(pk, orc, sc, lpf_mask=0, lpf_sigmas=None, use_kernels=True):
    tp = gpu_processor(pk, sc, use_kernels=use_kernels)
    lpf3 = None
    if lpf_sigmas is not None:
        tp.setLpfRbg(*lpf_sigmas)
        lpf3 = np.stack([orc.lpf(s) for s in lpf_sigmas])
    tp.execConverCorrectTiles(lpf_mask)
    g_clt = gpu_clt_cpu_layout(tp, sc)
    g = oracle_geom(orc, sc)
    kern = sc.kernels if (use_kernels and sc.kernels is not None) else None
    offs = sc.offsets if kern is not None else None
    o64 = orc.convert_frame(g, sc.images, sc.tasks, kernels=kern, offsets=offs, lpf_mask=lpf_mask, lpf3=lpf3)
    o32 = orc.convert_frame(g, sc.images, sc.tasks, kernels=kern, offsets=offs, lpf_mask=lpf_mask, lpf3=lpf3,
                            precision=32)
    tiles = task_tiles(sc)
    err = rel_rms(g_clt[:, tiles], o64[:, tiles])
    err32 = rel_rms(o32[:, tiles], o64[:, tiles])
    return tp, err, err32, o64


def test_convert_quad_bayer_kernels(pk, orc, quad_scene):
    log("case: convert quad 128x96 Bayer GR/BG, kernels, disparity 0..6")
    tp, err, err32, _ = _convert_case(pk, orc, quad_scene)
    log("  CLT rel RMS gpu=%.3g  fp32-oracle=%.3g" % (err, err32))
    assert err < CLT_TOL
    tp.close()


def test_convert_rggb_no_kernels_borders(pk, orc):
    log("case: convert quad 64x48 RGGB, no kernels, disparity 14 (windows cross the borders)")
    sc = pk.synth.make_scene(64, 48, seed=5, disparity=14.0, red_row=0, red_col=0, kernels=False)
    tp, err, err32, _ = _convert_case(pk, orc, sc)
    log("  CLT rel RMS gpu=%.3g  fp32-oracle=%.3g" % (err, err32))
    assert err < CLT_TOL
    tp.close()


def test_convert_lpf_mask(pk, orc, quad_scene):
    log("case: convert with the reference GPU LPF placement (lpf_mask 7, sigma 1.1/1.1/0.7)")
    tp, err, err32, _ = _convert_case(pk, orc, quad_scene, lpf_mask=7, lpf_sigmas=(1.1, 1.1, 0.7))
    log("  CLT rel RMS gpu=%.3g  fp32-oracle=%.3g" % (err, err32))
    assert err < CLT_TOL
    tp.close()


def test_convert_mono16_sparse(pk, orc):
    log("case: convert 16-camera mono 160x128, 30% sparse tasks")
    sc = pk.synth.make_scene(160, 128, num_cams=16, num_colors=1, seed=7, disparity=("field", 0.0, 4.0),
                             task_fraction=0.3)
    tp, err, err32, _ = _convert_case(pk, orc, sc)
    log("  CLT rel RMS gpu=%.3g  fp32-oracle=%.3g" % (err, err32))
    assert err < CLT_TOL
    tp.close()


def test_convert_skips_zero_tasks(pk, orc, quad_scene):
    log("case: task records with task word 0 are not converted")
    sc = quad_scene
    tasks = sc.tasks.copy()
    tasks[::3, 0] = 0
    tp = gpu_processor(pk, sc)
    tp.setTasks(tasks)
    tp.execConverCorrectTiles()
    g_clt = gpu_clt_cpu_layout(tp, sc)
    o = orc.convert_frame(oracle_geom(orc, sc), sc.images, tasks, kernels=sc.kernels, offsets=sc.offsets)
    live = tasks[:, 0] != 0
    tiles = ((tasks[:, 1] >> 16) * (sc.width // 8) + (tasks[:, 1] & 0xFFFF))[live]
    assert rel_rms(g_clt[:, tiles], o[:, tiles]) < CLT_TOL
    tp.close()


def test_imclt(pk, orc, quad_scene):
    sc = quad_scene
    for apply_lpf in (False, True):
        log("case: imclt quad 128x96, lpf=%s" % apply_lpf)
        tp = gpu_processor(pk, sc)
        tp.setCorrSigma(0.8)
        tp.execConverCorrectTiles()
        tp.execImcltRbg(apply_lpf)
        g_img = np.stack([tp.getRBG(c) for c in range(sc.num_cams)]).reshape(sc.num_cams, 3, sc.height + 8, sc.width + 8)
        g_clt = gpu_clt_cpu_layout(tp, sc).astype(np.float64)
        lpf = orc.lpf(0.8) if apply_lpf else None
        # same input: the GPU's own CLT through the oracle's inverse
        o_img = orc.imclt_frame(oracle_geom(orc, sc), g_clt, valid_map(sc), lpf64=lpf)
        err = rel_rms(g_img, o_img)
        # end to end against the FP64 convert + inverse
        o_clt = orc.convert_frame(oracle_geom(orc, sc), sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets)
        o_img2 = orc.imclt_frame(oracle_geom(orc, sc), o_clt, valid_map(sc), lpf64=lpf)
        err2 = rel_rms(g_img, o_img2)
        log("  image rel RMS same-input=%.3g end-to-end=%.3g" % (err, err2))
        assert err < IMG_TOL and err2 < IMG_TOL
        tp.close()


def test_imclt_sparse_tasks(pk, orc):
    log("case: imclt with a sparse task list (untouched tiles contribute nothing)")
    sc = pk.synth.make_scene(128, 96, seed=3, disparity=2.0, task_fraction=0.4)
    tp = gpu_processor(pk, sc)
    tp.execConverCorrectTiles()
    tp.execImcltRbg(False)
    g_img = np.stack([tp.getRBG(c) for c in range(sc.num_cams)]).reshape(sc.num_cams, 3, sc.height + 8, sc.width + 8)
    o_clt = orc.convert_frame(oracle_geom(orc, sc), sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets)
    o_img = orc.imclt_frame(oracle_geom(orc, sc), o_clt, valid_map(sc))
    err = rel_rms(g_img, o_img)
    log("  image rel RMS=%.3g" % err)
    assert err < IMG_TOL
    tp.close()


def _corr_compare(pk, orc, sc, fat_zero=0.0, pairs=None, strip=True, fused=False):
    tp = gpu_processor(pk, sc)
    tp.setCorrSigma(0.8)
    tp.setCorrParams(fat_zero=fat_zero)
    if pairs is not None:
        tp.setCorrPairs(pairs)
    if fused:
        tp.execConvertCorrelate2D()
    else:
        tp.execConverCorrectTiles()
        tp.execCorrelate2D()
    g_corr, g_combo, g_disp = tp.getCorr2D()
    g_clt = gpu_clt_cpu_layout(tp, sc).astype(np.float64)
    prm = orc.default_corr_params(fat_zero=fat_zero)
    o_corr, o_combo, o_disp = orc.corr_frame(oracle_geom(orc, sc), g_clt, sc.tasks, pairs=pairs, params=prm)
    o_clt = orc.convert_frame(oracle_geom(orc, sc), sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets)
    e_corr, e_combo, e_disp = orc.corr_frame(oracle_geom(orc, sc), o_clt, sc.tasks, pairs=pairs, params=prm)
    res = dict(corr=rel_rms(g_corr, o_corr), corr_e2e=rel_rms(g_corr, e_corr))
    if strip:
        assert np.array_equal(np.isnan(g_combo), np.isnan(o_combo)), "NaN pattern of the combined strip"
        res["combo"] = rel_rms(g_combo, o_combo)
        for name, k in (("int", 0), ("cm", 10), ("hor", 3), ("vert", 5)):   # cm: before the LMA
            for ref, tag in ((o_disp, ""), (e_disp, "_e2e")):
                both = np.isfinite(ref[:, k]) & np.isfinite(g_disp[:, k])
                assert np.array_equal(np.isfinite(ref[:, k]), np.isfinite(g_disp[:, k])), name + tag
                res[name + tag] = float(np.abs(g_disp[both, k] - ref[both, k]).max()) if both.any() else 0.0
        res["strength"] = rel_rms(g_disp[:, 11], o_disp[:, 11])
    tp.close()
    return res


def test_corr2d_quad(pk, orc, quad_scene):
    log("case: correlate2D quad 128x96, 6 pairs, fat_zero 0")
    r = _corr_compare(pk, orc, quad_scene)
    log("  " + ", ".join("%s=%.3g" % kv for kv in r.items()))
    assert r["corr"] < CLT_TOL and r["corr_e2e"] < CLT_TOL and r["combo"] < CLT_TOL
    assert r["int"] == 0 and r["cm"] < DISP_TOL and r["hor"] < DISP_TOL and r["vert"] < DISP_TOL
    assert r["cm_e2e"] < DISP_TOL and r["strength"] < CLT_TOL


def test_corr2d_fat_zero(pk, orc, quad_scene):
    log("case: correlate2D quad, fat_zero 0.05")
    r = _corr_compare(pk, orc, quad_scene, fat_zero=0.05)
    log("  " + ", ".join("%s=%.3g" % kv for kv in r.items()))
    assert r["corr"] < CLT_TOL and r["cm"] < DISP_TOL


@pytest.mark.parametrize("fat_zero", [0.0, 0.05])
def test_convert_corr2d_fused_against_oracle(pk, orc, quad_scene, fat_zero):
    log("case: convert + correlate2D in one kernel, quad 128x96, fat_zero %g" % fat_zero)
    r = _corr_compare(pk, orc, quad_scene, fat_zero=fat_zero, fused=True)
    log("  " + ", ".join("%s=%.3g" % kv for kv in r.items()))
    assert r["corr"] < CLT_TOL and r["corr_e2e"] < CLT_TOL and r["combo"] < CLT_TOL
    assert r["int"] == 0 and r["cm"] < DISP_TOL and r["hor"] < DISP_TOL and r["vert"] < DISP_TOL
    assert r["cm_e2e"] < DISP_TOL and r["strength"] < CLT_TOL


@pytest.mark.parametrize("fat_zero", [0.0, 0.05])
def test_convert_corr2d_fused_matches_two_calls(pk, quad_scene, fat_zero):
    log("case: execConvertCorrelate2D against execConverCorrectTiles + execCorrelate2D, every 3rd task "
        "word 0, fat_zero %g, twice in a row" % fat_zero)
    sc = quad_scene
    tasks = sc.tasks.copy()
    tasks[::3, 0] = 0
    out = []
    for fused in (False, True):
        tp = gpu_processor(pk, sc)
        tp.setTasks(tasks)
        tp.setCorrSigma(0.8)
        tp.setCorrParams(fat_zero=fat_zero)
        for _ in range(2):   # the second call writes the other correlation record set
            if fused:
                tp.execConvertCorrelate2D()
            else:
                tp.execConverCorrectTiles()
                tp.execCorrelate2D()
        out.append((gpu_clt_cpu_layout(tp, sc),) + tp.getCorr2D())
        tp.close()
    (c0, k0, s0, d0), (c1, k1, s1, d1) = out
    assert np.array_equal(c0, c1), "CLT"
    for name, a, b in (("corr", k0, k1), ("combo", s0, s1), ("disp", d0, d1)):
        assert np.array_equal(np.isnan(a), np.isnan(b)), name + " NaN pattern"
        fin = np.isfinite(a)
        err = float(np.abs(a[fin] - b[fin]).max()) if fin.any() else 0.0
        log("  %s max |fused - two calls| = %.3g" % (name, err))
        assert err <= 1e-6 * max(1.0, float(np.abs(a[fin]).max())), name
    live = tasks[:, 0] != 0
    assert np.isnan(k1[~live]).all()


def test_corr2d_mono16_all_pairs(pk, orc):
    log("case: correlate2D 16-camera mono 160x128, 30% tasks, 120 pairs")
    sc = pk.synth.make_scene(160, 128, num_cams=16, num_colors=1, seed=9, disparity=("field", 0.0, 4.0),
                             task_fraction=0.3)
    pairs = [(i, j, -1, 1) for i in range(16) for j in range(i + 1, 16)]
    r = _corr_compare(pk, orc, sc, pairs=pairs, strip=False)
    log("  " + ", ".join("%s=%.3g" % kv for kv in r.items()))
    assert r["corr"] < CLT_TOL and r["corr_e2e"] < CLT_TOL


def test_convert_corr2d_other_layouts_run_the_two_calls(pk):
    log("case: execConvertCorrelate2D outside the quad layout (3-camera Bayer, quad with 5 pairs) "
        "against execConverCorrectTiles + execCorrelate2D")
    cases = ((pk.synth.make_scene(96, 64, num_cams=3, seed=21, disparity=2.0), [(0, 1, 0, 1), (0, 2, 1, 1), (1, 2, -1, 1)]),
             (pk.synth.make_scene(96, 64, seed=22, disparity=2.0), [(0, 1, 0, 1), (2, 3, 0, 1), (0, 2, 1, 1),
                                                                    (1, 3, 1, 1), (0, 3, 2, 1)]))
    for sc, pairs in cases:
        out = []
        for fused in (False, True):
            tp = gpu_processor(pk, sc)
            tp.setCorrSigma(0.8)
            if pairs is not None:
                tp.setCorrPairs(pairs)
            if fused:
                tp.execConvertCorrelate2D()
            else:
                tp.execConverCorrectTiles()
                tp.execCorrelate2D()
            out.append((gpu_clt_cpu_layout(tp, sc),) + tp.getCorr2D())
            tp.close()
        for i, (a, b) in enumerate(zip(*out)):
            assert a.shape == b.shape and np.array_equal(a, b, equal_nan=True), (sc.num_cams, i)


def _tex_compare(pk, orc, sc, extra=None, tasks=None):
    tasks = sc.tasks if tasks is None else tasks
    tp = gpu_processor(pk, sc)
    tp.setTasks(tasks)
    tp.setCorrSigma(0.8)
    tp.execConverCorrectTiles()
    tp.execTextures(extra)
    g_tiles, g_stats = tp.getTextures()
    tp.execCombineTextures()
    g_img = tp.getCombinedTextures()
    g_clt = gpu_clt_cpu_layout(tp, sc).astype(np.float64)
    o_tiles, o_stats = orc.textures_frame(oracle_geom(orc, sc), g_clt, tasks, extra_disparity=extra)
    o_img = orc.combine_textures(oracle_geom(orc, sc), tasks, o_tiles)
    live = tasks[:, 0] != 0
    res = {}
    nc = sc.num_cams
    for name, sl in tex_planes(nc):
        res[name] = rel_rms(g_tiles[live, sl], o_tiles[live, sl])
    res["stats"] = rel_rms(g_stats[live], o_stats[live])
    res["image"] = rel_rms(g_img[:4], o_img[:4])
    tp.close()
    return res


def _assert_textures(r):
    """Textures on the same CLT (every pixel, no allowance): RGB, rms planes and the overlap-added
    image 1e-4 relative RMS, alpha 2e-4 and the port weights 2e-3 (helpers.TEX_TOL)."""
    for k, tol in TEX_TOL.items():
        assert r[k] < tol, k
    assert r["image"] < IMG_TOL
    assert r["stats"] < 1e-4


def test_textures_quad(pk, orc, quad_scene):
    log("case: textures quad 128x96, all ports")
    r = _tex_compare(pk, orc, quad_scene)
    log("  " + ", ".join("%s=%.3g" % kv for kv in r.items()))
    _assert_textures(r)


@pytest.mark.parametrize("cams", [2, 3, 6, 8])
def test_textures_camera_counts(pk, orc, cams):
    """Other port counts (cameras on a circle): the units of a tile take one pass of 16 lanes up to
    5 cameras and two passes from 6 on (the debayered tiles then get their own LDS)."""
    log("case: textures, %d cameras on a circle, 96x64" % cams)
    sc = pk.synth.make_scene(96, 64, num_cams=cams, seed=40 + cams, disparity=("field", 0.0, 4.0))
    r = _tex_compare(pk, orc, sc)
    log("  " + ", ".join("%s=%.3g" % kv for kv in r.items()))
    _assert_textures(r)


def test_textures_extra_disparity_port_masks(pk, orc, quad_scene):
    log("case: textures with extra disparity, partial port masks, forced-disparity bit")
    sc = quad_scene
    rng = np.random.default_rng(4)
    tasks = sc.tasks.copy()
    masks = np.array([0xF, 0x7, 0xB, 0x3, 0x1, 0x1000000 | 0xF], np.int32)
    tasks[:, 0] = masks[rng.integers(0, len(masks), len(tasks))]
    extra = rng.uniform(-1.5, 1.5, len(tasks)).astype(np.float32)
    r = _tex_compare(pk, orc, sc, extra=extra.astype(np.float64), tasks=tasks)
    log("  " + ", ".join("%s=%.3g" % kv for kv in r.items()))
    _assert_textures(r)


@pytest.mark.parametrize("phase,red", [("GR/BG", (0, 1)), ("RGGB", (0, 0))])
def test_config0_pipeline_disparity(pk, orc, phase, red):
    """4 x 512x512 Bayer, constant 3.0 px disparity, all tiles: convert + corr + imclt, in the
    reference's Bayer phase and in RGGB."""
    log("case: 512x512 quad frame %s, disparity 3.0 px, full pipeline" % phase)
    sc = pk.synth.make_scene(512, 512, seed=20261203, disparity=3.0, red_row=red[0], red_col=red[1])
    tp = gpu_processor(pk, sc)
    tp.setCorrSigma(0.8)
    tp.execConverCorrectTiles()
    tp.execCorrelate2D()
    tp.execImcltRbg(True)
    _, _, g_disp = tp.getCorr2D()
    g_img = np.stack([tp.getRBG(c) for c in range(4)]).reshape(4, 3, 520, 520)
    g = oracle_geom(orc, sc)
    o_clt = orc.convert_frame(g, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets)
    _, _, o_disp = orc.corr_frame(g, o_clt, sc.tasks)
    o_img = orc.imclt_frame(g, o_clt, valid_map(sc), lpf64=orc.lpf(0.8))
    g_clt = gpu_clt_cpu_layout(tp, sc)
    e_clt = rel_rms(g_clt, o_clt)
    e_img = rel_rms(g_img, o_img)
    fin = np.isfinite(o_disp[:, 10]) & np.isfinite(g_disp[:, 10])
    e_disp = float(np.abs(g_disp[fin, 10] - o_disp[fin, 10]).max())
    log("  CLT=%.3g image=%.3g cm-disparity max|d|=%.3g px over %d tiles" % (e_clt, e_img, e_disp, fin.sum()))
    assert e_clt < CLT_TOL and e_img < IMG_TOL and e_disp < DISP_TOL
    assert np.array_equal(np.isfinite(o_disp[:, 10]), np.isfinite(g_disp[:, 10]))
    tp.close()


def test_batch_matches_per_frame(pk, orc):
    """Three frames with uneven sparse task lists through one batched graph launch give the same
    results as each frame through the per-frame path; the autotuned shapes change nothing."""
    scenes = [pk.synth.make_scene(128, 96, seed=30 + f, disparity=("field", 0.0, 6.0)) for f in range(3)]
    rng = np.random.default_rng(5)
    keep = [rng.random(len(sc.tasks)) < frac for sc, frac in zip(scenes, (0.1, 1.0, 0.5))]
    bt = pk.GPUTileBatch(3, 128, 96)
    bt.setConvolutionKernels(scenes[0].kernels, offsets=scenes[0].offsets, force=True)
    bt.setCorrSigma(0.8)
    for f, sc in enumerate(scenes):
        bt.setBayerImagesFrame(f, sc.images)
        bt.setTasksFrame(f, sc.tasks[keep[f]])
    tune = bt.autotune()
    log("case: batch of 3 frames (10%%, 100%%, 50%% tasks), tune %s" % tune)
    bt.execBatch()
    bt.execBatch()   # graph replay
    for f, sc in enumerate(scenes):
        tp = pk.GPUTileProcessor(128, 96)
        tp.setConvolutionKernels(scenes[0].kernels, offsets=scenes[0].offsets, force=True)
        tp.setBayerImages(sc.images, force=True)
        tp.setTasks(sc.tasks[keep[f]])
        tp.setCorrSigma(0.8)
        tp.execConverCorrectTiles()
        tp.execCorrelate2D()
        tp.execTextures()
        tp.execImcltRbg(True)
        c1, s1, d1 = tp.getCorr2D()
        c2, s2, d2 = bt.getCorr2DFrame(f)
        t1, st1 = tp.getTextures()
        t2, st2 = bt.getTexturesFrame(f)
        for name, a, b in (("corr", c2, c1), ("combo", s2, s1), ("disp", d2, d1), ("tex", t2, t1), ("stats", st2, st1)):
            assert np.array_equal(np.isnan(a), np.isnan(b)), name
            m = ~np.isnan(b)
            assert np.abs(a[m] - b[m]).max() <= 1e-6 * max(1.0, np.abs(b[m]).max()), (f, name)
        for c in range(4):
            assert np.abs(bt.getRBGFrame(f, c) - tp.getRBG(c)).max() <= 1e-6 * max(1.0, np.abs(tp.getRBG(c)).max())
        tp.close()
    bt.close()


def test_batch_chunks_bit_identical(pk):
    """The pipelined batch over 2, 3 and 5 frame chunks (8 asked, capped at the frame count) gives
    bit-identical correlations, LMA records, textures and images to the one-chunk batch: chunks
    touch disjoint frames and task records.  One frame has no tasks (a chunk with no task
    launches); convert + imclt_rbg alone are checked the same way."""
    F = 5
    scenes = [pk.synth.make_scene(128, 96, seed=90 + f, disparity=("field", 0.0, 6.0)) for f in range(F)]
    rng = np.random.default_rng(11)
    keep = [rng.random(len(sc.tasks)) < frac for sc, frac in zip(scenes, (0.3, 0.0, 1.0, 0.6, 0.1))]
    bt = pk.GPUTileBatch(F, 128, 96)
    bt.setConvolutionKernels(scenes[0].kernels, offsets=scenes[0].offsets, force=True)
    bt.setCorrSigma(0.8)
    for f, sc in enumerate(scenes):
        bt.setBayerImagesFrame(f, sc.images)
        bt.setTasksFrame(f, sc.tasks[keep[f]])

    def results(stages):
        bt.execBatch(stages)
        bt.execBatch(stages)   # graph replay
        out = []
        for f in range(F):
            if stages & bt.CORR2D:
                out += list(bt.getCorr2DFrame(f))
            if stages & bt.TEXTURES:
                out += list(bt.getTexturesFrame(f))
            out += [bt.getRBGFrame(f, c) for c in range(4)]
        return out

    for stages in (15, bt.CONVERT | bt.IMCLT):
        bt.setBatchChunks(1)
        ref = results(stages)
        for chunks in (2, 3, 8):
            bt.setBatchChunks(chunks)
            log("case: batch of %d frames, stages %d, %d chunks (used %d)"
                % (F, stages, chunks, bt.getTune()["chunks"]))
            got = results(stages)
            assert len(got) == len(ref)
            for i, (a, b) in enumerate(zip(got, ref)):
                assert a.shape == b.shape and np.array_equal(a, b, equal_nan=True), (stages, chunks, i)
    assert bt.autotune()["chunks"] in (1, 2, 4)
    bt.close()


def test_batch_frame_with_idle_task(pk):
    """A frame whose task list holds a task with flags 0 cannot take the one-pass textures/imclt
    (its tile has no texture work), so the batch runs the two kernels for every frame; both
    frames still match the per-frame calls (NaN tile for the idle task)."""
    scenes = [pk.synth.make_scene(128, 96, seed=70 + f, disparity=("field", 0.0, 4.0)) for f in range(2)]
    tasks = [sc.tasks.copy() for sc in scenes]
    tasks[1][5, 0] = 0
    bt = pk.GPUTileBatch(2, 128, 96)
    bt.setConvolutionKernels(scenes[0].kernels, offsets=scenes[0].offsets, force=True)
    bt.setCorrSigma(0.8)
    for f, sc in enumerate(scenes):
        bt.setBayerImagesFrame(f, sc.images)
        bt.setTasksFrame(f, tasks[f])
    bt.execBatch()
    for f, sc in enumerate(scenes):
        tp = pk.GPUTileProcessor(128, 96)
        tp.setConvolutionKernels(scenes[0].kernels, offsets=scenes[0].offsets, force=True)
        tp.setBayerImages(sc.images, force=True)
        tp.setTasks(tasks[f])
        tp.setCorrSigma(0.8)
        tp.execConverCorrectTiles()
        tp.execTextures()
        tp.execImcltRbg(True)
        t1, s1 = tp.getTextures()
        t2, s2 = bt.getTexturesFrame(f)
        log("case: batch frame %d (idle task: %s)" % (f, f == 1))
        assert np.array_equal(np.isnan(t1), np.isnan(t2))
        assert np.array_equal(t1, t2, equal_nan=True) and np.array_equal(s1, s2, equal_nan=True)
        for c in range(4):
            assert np.abs(bt.getRBGFrame(f, c) - tp.getRBG(c)).max() <= 1e-6 * max(1.0, np.abs(tp.getRBG(c)).max())
        tp.close()
    bt.close()


def test_batch_follows_kernel_changes(pk):
    """A captured batch graph holds the kernel pointers of its capture: after setting or clearing
    the kernels the next clt_exec_batch must use the new ones.  Without kernels, then with them,
    then without again, each against the per-frame path under the same kernels."""
    scenes = [pk.synth.make_scene(128, 96, seed=60 + f, disparity=("field", 0.0, 4.0)) for f in range(2)]
    bt = pk.GPUTileBatch(2, 128, 96)
    bt.setCorrSigma(0.8)
    for f, sc in enumerate(scenes):
        bt.setBayerImagesFrame(f, sc.images)
        bt.setTasksFrame(f, sc.tasks)

    def per_frame(f, with_kernels):
        tp = pk.GPUTileProcessor(128, 96)
        if with_kernels:
            tp.setConvolutionKernels(scenes[0].kernels, offsets=scenes[0].offsets, force=True)
        tp.setBayerImages(scenes[f].images, force=True)
        tp.setTasks(scenes[f].tasks)
        tp.setCorrSigma(0.8)
        tp.execConverCorrectTiles()
        tp.execCorrelate2D()
        c, _, d = tp.getCorr2D()
        tp.close()
        return c, d

    for step, with_kernels in enumerate((False, True, False)):
        if with_kernels:
            bt.setConvolutionKernels(scenes[0].kernels, offsets=scenes[0].offsets, force=True)
        elif step:
            bt.clearConvolutionKernels()
        bt.execBatch()
        for f in range(2):
            c1, d1 = per_frame(f, with_kernels)
            c2, _, d2 = bt.getCorr2DFrame(f)
            err = float(np.abs(c2 - c1).max())
            log("case: batch step %d (kernels %s), frame %d: corr max |diff| %.3g" % (step, with_kernels, f, err))
            assert err <= 1e-6 * max(1.0, float(np.abs(c1).max()))
            assert np.array_equal(np.isnan(d1), np.isnan(d2))
    bt.close()


@pytest.mark.parametrize("overlap,sharp", [(True, False), (True, True), (False, False)])
def test_combine_rgba_tiles_modes(pk, orc, overlap, sharp):
    """combineRGBATiles (ID:3855-3991) on the device's own texture tiles against the oracle: the
    overlap-add, its sharp-alpha form and the mosaic of whole tiles, with a sparse task list and two
    tasks on one tile (the later one holds the position)."""
    log("case: combineRGBATiles overlap=%s sharp_alpha=%s, 128x96, 60%% of the tiles" % (overlap, sharp))
    sc = pk.synth.make_scene(128, 96, seed=27, disparity=("field", 0.0, 5.0), task_fraction=0.6)
    tasks = np.concatenate([sc.tasks, sc.tasks[3:4]])
    tasks[-1, 0] = 0x7
    tp = gpu_processor(pk, sc)
    tp.setTasks(tasks)
    tp.setCorrSigma(0.8)
    tp.execConverCorrectTiles()
    tp.execTextures()
    tiles, _ = tp.getTextures()
    tp.execCombineTextures(overlap=overlap, sharp_alpha=sharp)
    img = tp.getCombinedTextures()
    tp.close()
    ref = orc.combine_textures(oracle_geom(orc, sc), tasks, tiles.astype(np.float64), overlap=overlap, sharp_alpha=sharp)
    err = float(np.abs(img - ref).max() / np.abs(ref).max())
    log("  shape %s, max |diff| / max %.3g" % (img.shape, err))
    assert img.shape == ref.shape and err < 1e-6
