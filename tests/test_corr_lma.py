"""GPU parity of the per-tile correlation LMA chain (ImageDtt ID:2254-2336: Correlations2dLMA fit
over the combined pairs, the four one-direction fits and the foreground correction).

Same input: the oracle's FP64 chain runs on the GPU's own pair correlations and pre-LMA records, so
both sides take the same discrete decisions (sample selection, accepted steps) up to the order of
the FP64 sums; the final disparity, strength and LMA disparity agree to 1e-4 px on every tile.
End to end (oracle FP64 correlations from its own convert): every tile within 0.01 px, except the
tiles whose FP64 result is not determined at float32 precision (helpers.fp32_sensitive: perturbing
the FP64 CLT by one float32 ulp moves the oracle's own disparity past 0.01 px, e.g. a convexity test,
the rounding of the start point or an accepted step within rounding of its threshold); those are
found from the oracle alone and counted.
"""
import numpy as np
import pytest

from helpers import fp32_ill_conditioned, gpu_processor, log, oracle_geom

pytestmark = pytest.mark.gpu

DISP_TOL = 0.01
LAYERS = (("final", 1), ("strength", 2), ("poly", 8), ("poly_strength", 9))


def _same_input(orc, g_corr, g_disp, lma=None):
    o = orc.corr_lma_frame(g_corr, g_disp.astype(np.float64), lma=lma)
    res = {}
    for name, k in LAYERS:
        fo, fg = np.isfinite(o[:, k]), np.isfinite(g_disp[:, k])
        res[name + "_nan"] = int((fo != fg).sum())
        both = fo & fg
        res[name] = float(np.abs(g_disp[both, k] - o[both, k]).max()) if both.any() else 0.0
    res["fitted"] = int(np.isfinite(g_disp[:, 8]).sum())
    return res


def _end_to_end(g_disp, e_disp, ok):
    """Tiles past 0.01 px (and NaN mismatches) among the well-conditioned tiles `ok`."""
    res = {"ill_conditioned": int((~ok).sum())}
    for name, k in (("final", 1), ("poly", 8)):
        fo, fg = np.isfinite(e_disp[ok, k]), np.isfinite(g_disp[ok, k])
        both = fo & fg
        d = np.abs(g_disp[ok][both, k] - e_disp[ok][both, k])
        res[name + "_nan"] = int((fo != fg).sum())
        res[name + "_off"] = int((d > DISP_TOL).sum())
        res[name + "_median"] = float(np.median(d)) if both.any() else 0.0
    return res


def _run(pk, orc, sc, **lma):
    tp = gpu_processor(pk, sc)
    tp.setCorrSigma(0.8)
    if lma:
        tp.setCorrLmaParams(**lma)
    tp.execConverCorrectTiles()
    tp.execCorrelate2D()
    g_corr, _, g_disp = tp.getCorr2D()
    tp.close()
    return g_corr, g_disp


@pytest.mark.parametrize("name,size,kw", [
    ("field 0..6 px", (128, 96), dict(disparity=("field", 0.0, 6.0))),
    ("constant 3 px, residual 1 px", (128, 96), dict(disparity=3.0, target_disparity=2.0)),
    ("512x512 frame, field 0..20 px", (512, 512), dict(disparity=("field", 0.0, 20.0)))])
def test_corr_lma_quad(pk, orc, name, size, kw):
    log("case: per-tile LMA chain, quad %dx%d, %s" % (size[0], size[1], name))
    sc = pk.synth.make_scene(size[0], size[1], seed=21, **kw)
    g_corr, g_disp = _run(pk, orc, sc)
    r = _same_input(orc, g_corr, g_disp)
    log("  same input: " + ", ".join("%s=%.3g" % kv for kv in r.items()))
    g = oracle_geom(orc, sc)
    o_clt = orc.convert_frame(g, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets)
    ill, (_, _, e_disp) = fp32_ill_conditioned(orc, g, o_clt, sc.tasks)
    e = _end_to_end(g_disp, e_disp, ~ill)
    log("  end to end: " + ", ".join("%s=%.3g" % kv for kv in e.items()))
    assert r["fitted"] > 0.5 * len(sc.tasks)
    for nm, _ in LAYERS:
        assert r[nm + "_nan"] == 0 and r[nm] < 1e-4, nm
    assert e["final_off"] == 0 and e["poly_off"] == 0 and e["final_nan"] == 0 and e["poly_nan"] == 0
    assert e["ill_conditioned"] <= 0.02 * len(sc.tasks)   # small frames: many border tiles


def test_corr_lma_disabled_and_parameters(pk, orc):
    log("case: per-tile LMA off (records keep the centre of mass), then non-default parameters")
    sc = pk.synth.make_scene(128, 96, seed=22, disparity=("field", 0.0, 4.0))
    g_corr, g_disp = _run(pk, orc, sc, enable=0)
    assert np.isnan(g_disp[:, 8]).all()
    same = np.isfinite(g_disp[:, 10])
    assert np.array_equal(g_disp[same, 1], g_disp[same, 10]) and np.array_equal(g_disp[:, 2], g_disp[:, 11])
    kw = dict(mix_corr_poly=0, fo_correct=0, cnvx_hwnd_size=3, lma_adjust_wy=1, lma_adjust_wxy=0,
              ortho_vasw_pwr=1.0, lma_num_iter=8)
    g_corr, g_disp = _run(pk, orc, sc, **kw)
    o = orc.corr_lma_frame(g_corr, g_disp.astype(np.float64), lma=orc.default_corr_lma_params(**kw))
    both = np.isfinite(o[:, 8]) & np.isfinite(g_disp[:, 8])
    d = float(np.abs(g_disp[both, 8] - o[both, 8]).max())
    log("  non-default: %d fitted, poly max|d|=%.3g px" % (both.sum(), d))
    assert np.array_equal(np.isfinite(o[:, 8]), np.isfinite(g_disp[:, 8])) and d < 1e-4
    assert np.array_equal(g_disp[same, 1], g_disp[same, 10])   # mix off: CM unchanged


@pytest.mark.parametrize("name,kw", [
    ("Wyd adjusted (6 parameters, 4-core direction fits)", dict(lma_adjust_wy=1)),
    ("background correction, any direction, over-correction",
     dict(fo_far=1, fo_ortho=0, fo_overcorrection=0.5, fo_min_diff=0.05))])
def test_corr_lma_foreground_variants(pk, orc, name, kw):
    """The chain with the one-direction fits and the foreground correction on, under parameters the
    default run does not take: the NC = 4 fits (Wyd adjusted) and the far/any-direction/over-
    correction branches of foregroundCorrect (C2:1663-1750)."""
    log("case: per-tile LMA chain, %s" % name)
    sc = pk.synth.make_scene(128, 96, seed=23, disparity=("field", 0.0, 6.0))
    g_corr, g_disp = _run(pk, orc, sc, **kw)
    r = _same_input(orc, g_corr, g_disp, lma=orc.default_corr_lma_params(**kw))
    log("  same input: " + ", ".join("%s=%.3g" % kv for kv in r.items()))
    assert r["fitted"] > 0.5 * len(sc.tasks)
    for nm, _ in LAYERS:
        assert r[nm + "_nan"] == 0 and r[nm] < 1e-4, nm
    moved = int(np.sum(np.isfinite(g_disp[:, 1]) & (g_disp[:, 1] != g_disp[:, 8])))
    log("  tiles whose final disparity differs from the LMA's (foreground correction): %d" % moved)
