"""Correlation-peak loss and its hand-written adjoint.

CPU: the closed form of the zero-offset correlation (the quantity the loss is built from) equals
the centre of the oracle's 15x15 pair correlations.  GPU: the loss equals minus the sum of the
centres of correlate2D's output, and its gradient by the port centres matches FP64 central
differences of the same loss through the oracle's convert_direct (double-precision centres).
"""
import numpy as np
import pytest

from helpers import gpu_processor, log, oracle_geom

PAIRS = [(0, 1), (2, 3), (0, 2), (1, 3), (0, 3), (2, 1)]
COL_W = np.array([0.5, 0.2, 1.0]) / 1.7        # R, B, G weights (ImageDtt.java:1571-1581)


def centre_corr(D, lpf64):
    """C_p = 1/16 sum_k w_k sum_bins lpf <D1, D2> / (|D1| |D2|); D [4 cams][3 colours][4][64]."""
    out = []
    for a, b in PAIRS:
        s = 0.0
        for k in range(3):
            d1, d2 = D[a][k], D[b][k]
            f = (d1 * d2).sum(0) / np.sqrt((d1 * d1).sum(0) * (d2 * d2).sum(0))
            s += COL_W[k] * (lpf64 * f).sum()
        out.append(s / 16.0)
    return np.array(out)


def test_zero_offset_correlation_closed_form(orc, pk):
    sc = pk.synth.make_scene(64, 48, seed=3, disparity=2.0)
    g = oracle_geom(orc, sc)
    clt = orc.convert_frame(g, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets)
    corr, _, _ = orc.corr_frame(g, clt, sc.tasks)
    lpf = orc.lpf(0.8)
    for t in range(0, len(sc.tasks), 5):
        tile = (sc.tasks[t, 1] >> 16) * 8 + (sc.tasks[t, 1] & 0xFFFF)
        D = clt[:, tile].reshape(4, 3, 4, 64)
        assert np.abs(centre_corr(D, lpf) - corr[t].reshape(6, 15, 15)[:, 7, 7]).max() < 1e-8


def oracle_loss(orc, g, sc, t, xy, lpf):
    D = np.zeros((4, 3, 4, 64))
    for c in range(4):
        for k in range(3):
            D[c, k], _ = orc.convert_one(g, sc.images[c], k, xy[c][0], xy[c][1], kernels=sc.kernels[c],
                                         offsets=sc.offsets[c])
    return -centre_corr(D, lpf).sum()


@pytest.mark.gpu
def test_loss_matches_correlate2d_centres(pk):
    sc = pk.synth.make_scene(128, 96, seed=21, disparity=("field", 0.0, 6.0))
    tp = gpu_processor(pk, sc)
    tp.setCorrSigma(0.8)
    tp.execConverCorrectTiles()
    tp.execCorrelate2D()
    corr, _, _ = tp.getCorr2D()
    tp.execCorrAdjoint()
    adj = tp.getCorrAdjoint()
    centres = corr[:, :, 7, 7]
    err = np.abs(adj["c"] - centres).max()
    log("case: loss vs correlate2D centres, max |diff| = %.3g" % err)
    assert err < 1e-5
    assert np.abs(adj["loss"] + centres.sum(1)).max() < 1e-5
    assert (adj["kx"] > 0).mean() > 0.9 and (adj["ky"] > 0).mean() > 0.9
    tp.close()


@pytest.mark.gpu
def test_gradient_by_port_centres_matches_fp64_differences(pk, orc):
    sc = pk.synth.make_scene(128, 96, seed=22, disparity=("field", 0.0, 6.0))
    tp = gpu_processor(pk, sc)
    tp.setCorrSigma(0.8)
    tp.execCorrAdjoint()
    adj = tp.getCorrAdjoint()
    tp.close()
    g = oracle_geom(orc, sc)
    lpf = orc.lpf(0.8)
    rng = np.random.default_rng(4)
    h = 1e-4
    errs, scale = [], []
    for t in rng.choice(len(sc.tasks), 12, replace=False):
        xy = sc.tasks[t, 2:].view(np.float32).reshape(4, 2).astype(np.float64)
        l0 = oracle_loss(orc, g, sc, t, xy, lpf)
        assert abs(l0 - adj["loss"][t]) < 1e-4 * max(1.0, abs(l0))
        for c in range(4):
            for a in range(2):
                xp, xm = xy.copy(), xy.copy()
                xp[c, a] += h
                xm[c, a] -= h
                fd = (oracle_loss(orc, g, sc, t, xp, lpf) - oracle_loss(orc, g, sc, t, xm, lpf)) / (2 * h)
                errs.append(adj["dxy"][t, c, a] - fd)
                scale.append(abs(fd))
    errs = np.abs(np.array(errs))
    ref = max(np.max(scale), 1e-6)
    log("case: dl/dxy vs FP64 central differences: max |err| = %.3g (max |grad| %.3g)" % (errs.max(), ref))
    assert errs.max() < 2e-3 * ref
