"""GPU parity at the full sizes of the measured workloads, against the FP64 oracle on the same inputs.

* quad 2592x1936 (324 x 242 = 78,408 tiles), Bayer RGGB, dense tasks, disparity a smooth field
  0..20 px, random per-tile aberration kernels: convert_direct, correlate2D with the per-tile LMA
  chain, textures (RGBA + port weights + rms) and imclt_rbg (LPF), each output's error against
  the FP64 oracle printed; the asynchronous imclt_rbg at this size (about 0.3 ms on the side
  stream) followed at once by a convert.
* 16 x 640x512 monochrome cameras on a circle, a random 30 % of the tiles, all 120 pairs.
* a batch of 4 quad 2592x1936 frames, each with its own disparity field and a uniformly random
  5..100 % of the tiles: the batched outputs against the oracle frame by frame.
* the correlation-peak loss adjoint at 2592x1936 with extrinsics drawn N(0, 1e-3): gradients by
  the port centres, the 13 extrinsics and the disparity of 256 random tiles against FP64 central
  differences of the oracle's loss.

Bars: CLT and images 1e-4 relative RMS on every tile.  Correlations 1e-4 relative RMS and
disparities within 0.01 px (integer maxima equal) on every tile whose FP64 result is determined
at float32 precision, with no allowance; the others are found from the oracle alone
(helpers.fp32_sensitive: the FP64 CLT perturbed by one float32 ulp moves their FP64 outputs past
the bars; a task past a bar that the whole-frame draws did not mark is checked again with 24
draws of its own), counted and bounded by ILL_MAX.  The LMA chain on the GPU's own correlations
(same input) within 1e-4 px on every tile.  Textures: RGB and the rms planes 1e-4 relative RMS,
alpha 2e-4, the port-weight planes 2e-3; and in the quad frame every plane group within twice the
error of the oracle's own FP32 build on the same frame (the port weights are exponentials of squared
differences of nearly equal port values, and the combiner picks the two best ports by comparison:
a float32 restatement of the reference is itself 3e-4 .. 9e-4 relative RMS off in those planes).
"""
import os

import numpy as np
import pytest

from helpers import (confirm_ill_conditioned, fp32_ill_conditioned, gpu_clt_cpu_layout, gpu_processor, log,
                     oracle_geom, rel_rms, valid_map)

pytestmark = pytest.mark.gpu

W, H = 2592, 1936
TOL = 1e-4
DISP_TOL = 0.01
ILL_MAX = 0.01             # fraction of float32-ill-conditioned tasks (measured: 0.25 .. 0.57 %)
NTHREADS = min(16, os.cpu_count() or 1)
TEX_PLANES = (("rgb", slice(0, 3)), ("alpha", slice(3, 4)), ("weights", slice(4, 8)), ("rms", slice(8, 12)))
TEX_TOL = {"rgb": TOL, "alpha": 2e-4, "weights": 2e-3, "rms": TOL}


def _lma_off(g_disp, o_disp, ok):
    """Well-conditioned tasks whose final or LMA disparity is more than 0.01 px off."""
    bad = np.zeros(len(ok), bool)
    for k in (1, 8):
        both = np.isfinite(g_disp[:, k]) & np.isfinite(o_disp[:, k])
        bad[both] |= np.abs(g_disp[both, k] - o_disp[both, k]) > DISP_TOL
    return np.nonzero(bad & ok)[0]


def _disp_err(g, o, k):
    fo, fg = np.isfinite(o[:, k]), np.isfinite(g[:, k])
    both = fo & fg
    d = np.abs(g[both, k].astype(np.float64) - o[both, k]) if both.any() else np.zeros(1)
    return int((fo != fg).sum()), d


@pytest.fixture(scope="module")
def c1_scene(pk):
    log("scene: quad 2592x1936 RGGB, dense, disparity field 0..20 px, kernels")
    return pk.synth.make_scene(W, H, seed=20261204, disparity=("field", 0.0, 20.0), red_row=0, red_col=0)


@pytest.mark.timeout(600)
def test_c1_quad_full_frame(pk, orc, c1_scene):
    sc = c1_scene
    log("case: quad 2592x1936 RGGB frame, convert + correlate2D + LMA + textures + imclt vs FP64 oracle")
    tp = gpu_processor(pk, sc)
    tp.setCorrSigma(0.8)
    tp.execConverCorrectTiles()
    tp.execCorrelate2D()
    tp.execTextures()
    tp.execImcltRbg(True)
    g_clt = gpu_clt_cpu_layout(tp, sc)
    g_corr, g_combo, g_disp = tp.getCorr2D()
    g_tex, g_stats = tp.getTextures()
    g_img = np.stack([tp.getRBG(c) for c in range(4)]).reshape(4, 3, H + 8, W + 8)
    tp.close()
    g = oracle_geom(orc, sc)
    err = {}
    o_clt = orc.convert_frame(g, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets, nthreads=NTHREADS)
    err["clt"] = rel_rms(g_clt, o_clt)
    del g_clt
    o_img = orc.imclt_frame(g, o_clt, valid_map(sc), lpf64=orc.lpf(0.8), nthreads=NTHREADS)
    err["image"] = rel_rms(g_img, o_img)
    del o_img, g_img
    ill, (o_corr, o_combo, o_disp) = fp32_ill_conditioned(orc, g, o_clt, sc.tasks, nthreads=NTHREADS)
    ok = ~ill
    err["ill_conditioned"] = int(ill.sum())
    err["corr"] = rel_rms(g_corr[ok], o_corr[ok])
    err["corr_all"] = rel_rms(g_corr, o_corr)
    assert np.array_equal(np.isnan(g_combo[ok]), np.isnan(o_combo[ok])), "NaN pattern of the combined strips"
    err["combo"] = rel_rms(g_combo[ok], o_combo[ok])
    o_tex, o_stats = orc.textures_frame(g, o_clt, sc.tasks, nthreads=NTHREADS)
    # the oracle's FP32 build on the same frame: the expected single-precision error
    c32 = orc.convert_frame(g, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets, nthreads=NTHREADS,
                            precision=32)
    t32, _ = orc.textures_frame(g, c32, sc.tasks, nthreads=NTHREADS, precision=32)
    del c32
    for name, sl in TEX_PLANES:
        err["tex_" + name] = rel_rms(g_tex[:, sl], o_tex[:, sl])
        err["tex_" + name + "_fp32"] = rel_rms(t32[:, sl], o_tex[:, sl])
        rms = np.sqrt(np.mean(o_tex[:, sl] ** 2))
        err["tex_" + name + "_max"] = float(np.abs(g_tex[:, sl] - o_tex[:, sl]).max() / rms)
    err["tex_stats"] = rel_rms(g_stats, o_stats)
    del o_tex, t32
    # disparities on the well-conditioned tiles: integer maxima equal, the other layers within 0.01 px
    nan_int, d_int = _disp_err(g_disp[ok], o_disp[ok], 0)
    err["disp_int"] = float(d_int.max())
    for name, k in (("cm", 10), ("hor", 3), ("vert", 5)):
        nn, d = _disp_err(g_disp[ok], o_disp[ok], k)
        err["disp_" + name], err["disp_" + name + "_nan"] = float(d.max()), nn
    # the LMA chain: same input (the oracle's chain on the GPU's correlations) on every tile, and
    # end to end on the well-conditioned tiles
    same = orc.corr_lma_frame(g_corr, g_disp.astype(np.float64), nthreads=NTHREADS)
    for name, k in (("final", 1), ("poly", 8)):
        nn, d = _disp_err(g_disp, same, k)
        err["lma_same_" + name], err["lma_same_" + name + "_nan"] = float(d.max()), nn
        nn, d = _disp_err(g_disp[ok], o_disp[ok], k)
        err["lma_e2e_" + name], err["lma_e2e_" + name + "_nan"] = float(d.max()), nn
        err["lma_e2e_" + name + "_off"] = int((d > DISP_TOL).sum())
        err["lma_e2e_" + name + "_median"] = float(np.median(d))
    # tiles past 0.01 px that the whole-frame draws left unmarked: 24 draws of their own
    off = _lma_off(g_disp, o_disp, ok)
    err["lma_recheck"] = len(off)
    off = np.setdiff1d(off, confirm_ill_conditioned(orc, g, o_clt, sc.tasks, off, nthreads=NTHREADS))
    del o_clt
    err["lma_e2e_off"] = len(off)
    # all tiles, for the record
    nn, d = _disp_err(g_disp, o_disp, 1)
    err["final_all_off"] = int((d > DISP_TOL).sum()) + nn
    log("  FP64 errors: " + ", ".join("%s=%.3g" % kv for kv in err.items()))
    assert err["clt"] < TOL and err["image"] < TOL and err["corr"] < TOL and err["combo"] < TOL
    assert err["ill_conditioned"] <= ILL_MAX * len(sc.tasks)
    assert nan_int == 0 and err["disp_int"] == 0
    for name in ("cm", "hor", "vert"):
        assert err["disp_" + name + "_nan"] == 0 and err["disp_" + name] < DISP_TOL, name
    for name in ("final", "poly"):
        assert err["lma_same_" + name + "_nan"] == 0 and err["lma_same_" + name] < TOL, name
        assert err["lma_e2e_" + name + "_nan"] == 0, name
    assert err["lma_e2e_off"] == 0
    for name, _ in TEX_PLANES:
        assert err["tex_" + name] < TEX_TOL[name] and err["tex_" + name] <= 2 * err["tex_" + name + "_fp32"], name
    assert err["tex_stats"] < TOL


@pytest.mark.timeout(300)
def test_c1_async_imclt_then_convert(pk, c1_scene):
    """At this size imclt_rbg runs about 0.3 ms on its side stream: a convert issued right after it
    (async mode, no host synchronisation) must wait for it (the CLT it reads is rewritten)."""
    sc = c1_scene
    log("case: async imclt_rbg at 2592x1936 followed at once by a convert with another LPF")

    def run(asynchronous):
        tp = gpu_processor(pk, sc)
        tp.setCorrSigma(0.8)
        tp.setLpfRbg(0.7, 0.7, 0.7)
        tp.setAsync(asynchronous)
        tp.execConverCorrectTiles(0)
        tp.execImcltRbg(True)
        tp.execConverCorrectTiles(7)
        img = tp.getRBG(1)
        clt = tp.getCLT(2)
        tp.close()
        return img, clt

    i0, c0 = run(False)
    i1, c1 = run(True)
    same_img = np.array_equal(i0.view(np.int32), i1.view(np.int32))
    same_clt = np.array_equal(c0.view(np.int32), c1.view(np.int32))
    log("  bit-identical: images=%s, second CLT=%s" % (same_img, same_clt))
    assert same_img and same_clt


@pytest.mark.timeout(600)
def test_c1_async_corr2d_while_lma_runs(pk, c1_scene):
    """Async frames at this size: each convert writes the other CLT buffer on its own stream while
    the previous frame's correlate2D, one-pass textures/imclt and LMA (about 0.4 ms on its own
    stream) may still run; correlate2D issued before the previous LMA ends writes the second pair
    of correlation records, and a third frame goes back to the first buffers after waiting for
    their readers.  The last frame's correlations, disparities, textures and images must be
    bit-identical to the same frame run synchronously."""
    sc = c1_scene
    log("case: async frames x2 / x3 at 2592x1936 (convert, correlate2D + LMA, textures + imclt)")

    def run(asynchronous, lpf_masks):
        tp = gpu_processor(pk, sc)
        tp.setCorrSigma(0.8)
        tp.setLpfRbg(0.7, 0.7, 0.7)
        tp.setTexParams(keep_weights=0)
        tp.setAsync(asynchronous)
        for m in lpf_masks:
            tp.execConverCorrectTiles(m)
            tp.execCorrelate2D()
            tp.execTexturesImclt(True)
        out = list(tp.getCorr2D()) + list(tp.getTextures()) + [tp.getRBG(1), tp.getCLT(2)]
        tp.close()
        return out

    for masks in ((0, 7), (0, 7, 0)):
        ref = run(False, masks[-1:])
        got = run(True, masks)
        same = [np.array_equal(a.view(np.int32), b.view(np.int32)) for a, b in zip(got, ref)]
        log("  %d frames: bit-identical corr/combo/disparity/textures/stats/image/CLT = %s" % (len(masks), same))
        assert all(same)


@pytest.mark.timeout(300)
def test_c2_mono16_full_size(pk, orc):
    log("case: 16 x 640x512 mono, 30% tasks, 120 pairs: convert + correlate2D + imclt vs FP64 oracle")
    sc = pk.synth.make_scene(640, 512, num_cams=16, num_colors=1, seed=20261206, disparity=("field", 0.0, 6.0),
                             task_fraction=0.3)
    pairs = [(i, j, -1, 1) for i in range(16) for j in range(i + 1, 16)]
    tp = gpu_processor(pk, sc)
    tp.setCorrSigma(0.8)
    tp.setCorrPairs(pairs)
    tp.execConverCorrectTiles()
    tp.execCorrelate2D()
    tp.execImcltRbg(True)
    g_clt = gpu_clt_cpu_layout(tp, sc)
    g_corr, _, _ = tp.getCorr2D()
    g_img = np.stack([tp.getRBG(c) for c in range(16)]).reshape(16, 1, 520, 648)
    tp.close()
    g = oracle_geom(orc, sc)
    o_clt = orc.convert_frame(g, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets, nthreads=NTHREADS)
    o_corr, _, _ = orc.corr_frame(g, o_clt, sc.tasks, pairs=pairs, nthreads=NTHREADS)
    o_img = orc.imclt_frame(g, o_clt, valid_map(sc), lpf64=orc.lpf(0.8), nthreads=NTHREADS)
    live = sc.tasks[:, 0] != 0
    tiles = ((sc.tasks[:, 1] >> 16) * 80 + (sc.tasks[:, 1] & 0xFFFF))[live]
    err = {"tasks": len(sc.tasks), "clt": rel_rms(g_clt[:, tiles], o_clt[:, tiles]), "corr": rel_rms(g_corr, o_corr),
           "image": rel_rms(g_img, o_img)}
    # the 15x15 peak of every pair: the GPU's argmax is the oracle's, or a tie in the oracle's
    # FP64 values (within 1e-5 of the peak value)
    gc = g_corr.reshape(len(g_corr), 120, 225)
    oc = o_corr.reshape(len(o_corr), 120, 225)
    ga, oa = gc.argmax(-1), oc.argmax(-1)
    err["argmax_diff"] = int((ga != oa).sum())
    at_g = np.take_along_axis(oc, ga[..., None], -1)[..., 0]
    gap = np.abs(oc.max(-1) - at_g) / np.abs(oc).max(-1)
    err["argmax_gap"] = float(gap.max())
    log("  FP64 errors: " + ", ".join("%s=%.3g" % kv for kv in err.items()))
    assert err["clt"] < TOL and err["corr"] < TOL and err["image"] < TOL
    assert err["argmax_gap"] < 1e-5


@pytest.mark.timeout(900)
def test_c3_batch_full_frames_vs_oracle(pk, orc):
    """4 quad 2592x1936 frames with independent disparity fields and task coverage through one
    batched launch; every frame's correlations, disparities, textures and images against the
    oracle's FP64 pipeline of that frame."""
    F = 4
    rng = np.random.default_rng(20261207)
    cover = rng.uniform(0.05, 1.0, F)
    kx, ky = W // 16 + 2, H // 16 + 2
    kset = pk.synth.clt_kernels(4, 3, ky, kx, np.random.default_rng(20261207 + 7))
    bt = pk.GPUTileBatch(F, W, H)
    bt.setConvolutionKernels(kset[0], offsets=kset[1], force=True)
    bt.setCorrSigma(0.8)
    scenes = []
    for f in range(F):
        sc = pk.synth.make_scene(W, H, seed=20261207 + 1000 * f, disparity=("field", 0.0, 20.0),
                                 task_fraction=float(cover[f]), kernels=kset)
        bt.setBayerImagesFrame(f, sc.images)
        bt.setTasksFrame(f, sc.tasks)
        scenes.append(sc)
        log("  frame %d: %d tasks (%.0f %%)" % (f, len(sc.tasks), 100 * len(sc.tasks) / (kx - 2) / (ky - 2)))
    tune = bt.autotune()
    log("case: batch of %d full frames, %d tasks, tune %s" % (F, sum(len(s.tasks) for s in scenes), tune))
    bt.execBatch()
    bt.execBatch()   # graph replay
    worst = {}
    for f, sc in enumerate(scenes):
        g = oracle_geom(orc, sc)
        o_clt = orc.convert_frame(g, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets, nthreads=NTHREADS)
        ill, (o_corr, o_combo, o_disp) = fp32_ill_conditioned(orc, g, o_clt, sc.tasks, nthreads=NTHREADS)
        ok = ~ill
        o_tex, o_stats = orc.textures_frame(g, o_clt, sc.tasks, nthreads=NTHREADS)
        o_img = orc.imclt_frame(g, o_clt, valid_map(sc), lpf64=orc.lpf(0.8), nthreads=NTHREADS)
        c, s, d = bt.getCorr2DFrame(f)
        t, st = bt.getTexturesFrame(f)
        img = np.stack([bt.getRBGFrame(f, cam) for cam in range(4)]).reshape(4, 3, H + 8, W + 8)
        same = orc.corr_lma_frame(c, d.astype(np.float64), nthreads=NTHREADS)
        e = {"ill": float(ill.mean()), "corr": rel_rms(c[ok], o_corr[ok]), "combo": rel_rms(s[ok], o_combo[ok]),
             "image": rel_rms(img, o_img), "stats": rel_rms(st, o_stats)}
        for name, sl in TEX_PLANES:
            e["tex_" + name] = rel_rms(t[:, sl], o_tex[:, sl])
        assert np.array_equal(np.isnan(s[ok]), np.isnan(o_combo[ok])), f
        nn, dd = _disp_err(d[ok], o_disp[ok], 0)
        e["int"] = float(dd.max()) + nn
        nn, dd = _disp_err(d[ok], o_disp[ok], 10)
        e["cm"], e["cm_nan"] = float(dd.max()), nn
        nn, dd = _disp_err(d, same, 1)
        e["lma_same"], e["lma_same_nan"] = float(dd.max()), nn
        nn, dd = _disp_err(d[ok], o_disp[ok], 1)
        off = _lma_off(d, o_disp, ok)
        e["lma_recheck"] = len(off)
        off = np.setdiff1d(off, confirm_ill_conditioned(orc, g, o_clt, sc.tasks, off, nthreads=NTHREADS))
        del o_clt
        e["lma_e2e_off"] = len(off) + nn
        log("  frame %d: " % f + ", ".join("%s=%.3g" % kv for kv in e.items()))
        for k, v in e.items():
            worst[k] = max(worst.get(k, 0.0), v)
    bt.close()
    assert worst["corr"] < TOL and worst["combo"] < TOL and worst["image"] < TOL
    assert worst["int"] == 0 and worst["cm"] < DISP_TOL and worst["cm_nan"] == 0
    assert worst["lma_same"] < TOL and worst["lma_same_nan"] == 0 and worst["lma_e2e_off"] == 0
    assert worst["ill"] <= ILL_MAX
    for name, _ in TEX_PLANES:
        assert worst["tex_" + name] < TEX_TOL[name], name
    assert worst["stats"] < TOL


# ----------------------------------------------------------------------------------- adjoint
PAIRS = [(0, 1), (2, 3), (0, 2), (1, 3), (0, 3), (2, 1)]
COL_W = np.array([0.5, 0.2, 1.0]) / 1.7        # R, B, G weights (ImageDtt.java:1571-1581)


def _tile_losses(orc, g, sc, tiles, xy, lpf):
    """Loss of each of the given tiles with port centres xy [n][4][2] (double): minus the sum of
    the zero-offset pair correlations of the oracle's FP64 convert (pinned against the centres of
    the oracle's 15x15 correlations in test_adjoint.py)."""
    n = len(tiles)
    clt = np.zeros((4, n, 3, 4, 64))
    # convert_one takes the centres in double (a task record holds them as float32)
    for i in range(n):
        for c in range(4):
            for k in range(3):
                clt[c, i, k], _ = orc.convert_one(g, sc.images[c], k, xy[i, c, 0], xy[i, c, 1], kernels=sc.kernels[c],
                                                  offsets=sc.offsets[c])
    loss = np.zeros(n)
    for a, b in PAIRS:
        for k in range(3):
            d1, d2 = clt[a, :, k], clt[b, :, k]
            f = (d1 * d2).sum(1) / np.sqrt((d1 * d1).sum(1) * (d2 * d2).sum(1))
            loss -= COL_W[k] * (lpf[None, :] * f).sum(1) / 16.0
    return loss


@pytest.mark.timeout(900)
def test_c4_adjoint_full_frame_fd(pk, orc, c1_scene):
    """Adjoint on the 2592x1936 frame with extrinsics drawn N(0, 1e-3): the device gradients by
    the port centres (per task), by the 13 extrinsics and by each task's disparity against FP64
    central differences (h = 1e-4 px of port motion) of the oracle's loss on 256 random tiles.
    Tiles whose forward and backward differences disagree (a window or kernel-tile rounding step
    inside +-h) are left out and counted."""
    sc = c1_scene
    rng = np.random.default_rng(20261204 + 13)
    ext = rng.normal(0, 1e-3, 13)
    tx_all, ty_all = np.meshgrid(np.arange(W // 8), np.arange(H // 8))
    inner = ((tx_all > 2) & (tx_all < W // 8 - 3) & (ty_all > 2) & (ty_all < H // 8 - 3)).ravel()
    tiles = np.sort(rng.choice(np.nonzero(inner)[0], 256, replace=False))
    tx, ty = tiles % (W // 8), tiles // (W // 8)
    d0 = sc.disparity[ty, tx].astype(np.float32)
    tp = pk.GPUTileProcessor(W, H, 4, 3, red_row=0, red_col=0)
    tp.setConvolutionKernels(sc.kernels, offsets=sc.offsets, force=True)
    tp.setBayerImages(sc.images, force=True)
    tp.setCorrSigma(0.8)
    pg = pk.GPUTileProcessor.defaultPortGeometry(W, H, extrinsic=list(ext))
    tp.setPortGeometry(pg)
    tp.getTileSubcamOffsets(np.full(256, 0xFF, np.int32), tx, ty, d0)
    tp.execCorrAdjoint()
    adj = tp.getCorrAdjoint()
    ge, gd, loss_gpu = tp.getGeometryGradient()
    tp.close()
    g = orc.make_geom(W, H, red_row=0, red_col=0)
    lpf = orc.lpf(0.8)
    opg = orc.default_port_geometry(W, H, extrinsic=list(ext))
    xy = orc.port_coords(opg, tx * 8 + 4.0, ty * 8 + 4.0, d0.astype(np.float64))
    l0 = _tile_losses(orc, g, sc, tiles, xy, lpf)
    log("case: adjoint 2592x1936, 256 tiles, extrinsics N(0,1e-3): loss gpu %.6f oracle %.6f" % (loss_gpu, l0.sum()))
    assert abs(loss_gpu - l0.sum()) < 1e-4 * abs(l0.sum())
    h = 1e-4
    # gradient by the port centres, all tiles at once (each tile's loss depends on its own centres)
    fd = np.zeros((256, 4, 2))
    bad = np.zeros(256, bool)
    for c in range(4):
        for a in range(2):
            xp, xm = xy.copy(), xy.copy()
            xp[:, c, a] += h
            xm[:, c, a] -= h
            lp, lm = _tile_losses(orc, g, sc, tiles, xp, lpf), _tile_losses(orc, g, sc, tiles, xm, lpf)
            fwd, bwd = (lp - l0) / h, (l0 - lm) / h
            fd[:, c, a] = (lp - lm) / (2 * h)
            bad |= np.abs(fwd - bwd) > 0.05 * np.abs(fd[:, c, a]) + 1e-3
    ok = ~bad
    scale = np.abs(fd[ok]).max()
    e_xy = np.abs(adj["dxy"][ok] - fd[ok]).max() / scale
    # gradient by the disparity of each tile
    d64 = d0.astype(np.float64)   # the steps in double: d0 + h in float32 would round h by up to 1 %
    dp = orc.port_coords(opg, tx * 8 + 4.0, ty * 8 + 4.0, d64 + h)
    dm = orc.port_coords(opg, tx * 8 + 4.0, ty * 8 + 4.0, d64 - h)
    fd_d = (_tile_losses(orc, g, sc, tiles, dp, lpf) - _tile_losses(orc, g, sc, tiles, dm, lpf)) / (2 * h)
    e_d = np.abs(gd[ok] - fd_d[ok]).max() / np.abs(fd_d[ok]).max()
    # gradient by the extrinsics, over the tiles kept above (1e-4 px of port motion: h / f_pix)
    he = h / (4.5 / 0.0022)
    keep_t = tiles[ok]
    fd_e = np.zeros(13)
    for j in range(13):
        ep, em = ext.copy(), ext.copy()
        ep[j] += he
        em[j] -= he
        xp = orc.port_coords(orc.default_port_geometry(W, H, extrinsic=list(ep)), tx[ok] * 8 + 4.0, ty[ok] * 8 + 4.0,
                             d64[ok])
        xm = orc.port_coords(orc.default_port_geometry(W, H, extrinsic=list(em)), tx[ok] * 8 + 4.0, ty[ok] * 8 + 4.0,
                             d64[ok])
        fd_e[j] = (_tile_losses(orc, g, sc, keep_t, xp, lpf).sum() - _tile_losses(orc, g, sc, keep_t, xm, lpf).sum()) / (2 * he)
    # the device extrinsic gradient sums every task: subtract the left-out tiles' share by the chain rule
    ge_kept = ge.copy()
    if bad.any():
        tp2 = pk.GPUTileProcessor(W, H, 4, 3, red_row=0, red_col=0)
        tp2.setConvolutionKernels(sc.kernels, offsets=sc.offsets, force=True)
        tp2.setBayerImages(sc.images, force=True)
        tp2.setCorrSigma(0.8)
        tp2.setPortGeometry(pg)
        tp2.getTileSubcamOffsets(np.full(ok.sum(), 0xFF, np.int32), tx[ok], ty[ok], d0[ok])
        ge_kept, _, _ = tp2.getGeometryGradient()
        tp2.close()
    e_e = np.abs(ge_kept - fd_e).max() / np.abs(fd_e).max()
    log("  tiles left out (rounding step within +-h): %d of 256" % bad.sum())
    log("  dl/dxy max rel err %.3g (max |grad| %.3g); dl/dd %.3g (max |grad| %.3g); dl/de %.3g (max |grad| %.3g)" % (
        e_xy, scale, e_d, np.abs(fd_d[ok]).max(), e_e, np.abs(fd_e).max()))
    assert bad.sum() <= 8
    assert e_xy < 2e-3 and e_d < 2e-3 and e_e < 2e-3
