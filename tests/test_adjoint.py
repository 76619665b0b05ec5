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
    ok = ~np.isnan(adj["loss"])   # tasks with every window inside the frame
    xy = sc.tasks[:, 2:].view(np.float32).reshape(-1, 4, 2)
    inner = np.all((xy[..., 0] >= 9) & (xy[..., 0] <= 128 - 9) & (xy[..., 1] >= 9) & (xy[..., 1] <= 96 - 9), axis=1)
    assert np.all(ok[inner]) and ok.sum() > 0.5 * len(ok)
    assert np.all(adj["dxy"][~ok] == 0)
    err = np.abs(adj["c"][ok] - centres[ok]).max()
    log("case: loss vs correlate2D centres, max |diff| = %.3g over %d tasks" % (err, ok.sum()))
    assert err < 1e-5
    assert np.abs(adj["loss"][ok] + centres[ok].sum(1)).max() < 1e-5
    assert (adj["kx"][ok] > 0).mean() > 0.9 and (adj["ky"][ok] > 0).mean() > 0.9
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
    for t in rng.choice(np.nonzero(~np.isnan(adj["loss"]))[0], 12, replace=False):
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


def _geometry_processor(pk, sc, extrinsic, tiles, disparity):
    tp = pk.GPUTileProcessor(sc.width, sc.height, 4, 3, kernels=False)
    tp.setBayerImages(sc.images, force=True)
    tp.setCorrSigma(0.8)
    tp.setPortGeometry(pk.GPUTileProcessor.defaultPortGeometry(sc.width, sc.height, extrinsic=list(extrinsic)))
    tp.getTileSubcamOffsets(np.full(len(tiles), 0xFF, np.int32), sc.tx[tiles], sc.ty[tiles], disparity)
    return tp


def _oracle_total_loss(orc, sc, extrinsic, tiles, disparity, g, lpf):
    pg = orc.default_port_geometry(sc.width, sc.height, extrinsic=list(extrinsic))
    xy = orc.port_coords(pg, sc.tx[tiles] * 8 + 4.0, sc.ty[tiles] * 8 + 4.0, np.asarray(disparity, np.float64))
    total = 0.0
    for t in range(len(tiles)):
        D = np.zeros((4, 3, 4, 64))
        for c in range(4):
            for k in range(3):
                D[c, k], _ = orc.convert_one(g, sc.images[c], k, xy[t, c, 0], xy[t, c, 1])
        total -= centre_corr(D, lpf).sum()
    return total


@pytest.mark.gpu
def test_geometry_gradient_matches_fp64_differences(pk, orc):
    """dl/d(13 extrinsics) and dl/d(disparity) through the device adjoint against FP64 central
    differences of the oracle's loss (port geometry in double, convert at double centres)."""
    rng = np.random.default_rng(8)
    sc = pk.synth.make_geometry_scene(256, 192, rng.normal(0, 1e-3, 13), seed=8)
    tiles = rng.choice(len(sc.tx), 64, replace=False)
    tiles = tiles[(sc.tx[tiles] > 1) & (sc.tx[tiles] < 30) & (sc.ty[tiles] > 1) & (sc.ty[tiles] < 22)]
    e0 = sc.extrinsic + rng.normal(0, 3e-4, 13)
    d0 = (sc.disparity_tiles.ravel()[tiles] + rng.normal(0, 0.1, len(tiles))).astype(np.float32)
    tp = _geometry_processor(pk, sc, e0, tiles, d0)
    ge, gd, loss = tp.getGeometryGradient()
    tp.close()
    g = orc.make_geom(sc.width, sc.height, kernels=False)
    lpf = orc.lpf(0.8)
    d64 = d0.astype(np.float64)
    l0 = _oracle_total_loss(orc, sc, e0, tiles, d64, g, lpf)
    log("case: geometry adjoint, %d tiles, loss gpu %.6f oracle %.6f" % (len(tiles), loss, l0))
    assert abs(loss - l0) < 1e-4 * abs(l0)
    he = 2e-7
    fd_e = np.zeros(13)
    for j in range(13):
        ep, em = e0.copy(), e0.copy()
        ep[j] += he
        em[j] -= he
        fd_e[j] = (_oracle_total_loss(orc, sc, ep, tiles, d64, g, lpf) -
                   _oracle_total_loss(orc, sc, em, tiles, d64, g, lpf)) / (2 * he)
    err_e = np.abs(ge - fd_e).max() / np.abs(fd_e).max()
    hd = 1e-4
    fd_d, an_d = [], []
    for t in range(0, len(tiles), 7):
        dp, dm = d64.copy(), d64.copy()
        dp[t] += hd
        dm[t] -= hd
        fd_d.append((_oracle_total_loss(orc, sc, e0, tiles[t:t + 1], dp[t:t + 1], g, lpf) -
                     _oracle_total_loss(orc, sc, e0, tiles[t:t + 1], dm[t:t + 1], g, lpf)) / (2 * hd))
        an_d.append(gd[t])
    fd_d, an_d = np.array(fd_d), np.array(an_d)
    err_d = np.abs(an_d - fd_d).max() / np.abs(fd_d).max()
    log("  dl/de: max rel err %.3g (|grad| %.3g); dl/dd: max rel err %.3g" % (err_e, np.abs(fd_e).max(), err_d))
    assert err_e < 2e-3 and err_d < 2e-3


@pytest.mark.gpu
def test_lma_reduces_misalignment(pk, orc):
    """Scene rendered with extrinsics drawn N(0, 3e-4) (relative port offsets ~1.2 px, inside the
    capture range of a zero-offset correlation): 10 LMA iterations from zero extrinsics, with the
    disparities as parameters too, shrink the error of the relative port offsets (what the
    correlations see) more than 5 times, and the loss never increases."""
    rng = np.random.default_rng(9)
    e_true = rng.normal(0, 3e-4, 13)
    sc = pk.synth.make_geometry_scene(512, 384, e_true, seed=9)
    inner = (sc.tx > 1) & (sc.tx < sc.width // 8 - 2) & (sc.ty > 1) & (sc.ty < sc.height // 8 - 2)
    tiles = np.nonzero(inner)[0]
    d_true = sc.disparity_tiles.ravel()[tiles]
    d0 = (d_true + rng.normal(0, 0.05, len(tiles))).astype(np.float32)
    tp = _geometry_processor(pk, sc, np.zeros(13), tiles, d0)
    ext, disp, hist = tp.lmaExtrinsics(10)
    tp.close()

    def rel_err(e, d):
        px, py = sc.tx[tiles] * 8 + 4.0, sc.ty[tiles] * 8 + 4.0
        a = orc.port_coords(orc.default_port_geometry(sc.width, sc.height, extrinsic=list(e)), px, py, d)
        b = orc.port_coords(orc.default_port_geometry(sc.width, sc.height, extrinsic=list(e_true)), px, py, d_true)
        rel = lambda x: np.stack([x[:, i] - x[:, j] for i, j in PAIRS], 1)
        return float(np.sqrt(np.mean((rel(a) - rel(b)) ** 2)))

    before, after = rel_err(np.zeros(13), d0.astype(np.float64)), rel_err(ext, disp.astype(np.float64))
    log("case: LMA %d iterations: loss %.3f -> %.3f, relative port offset error %.4f -> %.4f px"
        % (len(hist) - 1, hist[0], hist[-1], before, after))
    log("  history %s" % np.array2string(hist, precision=2, max_line_width=200))
    log("  extrinsics true %s" % np.array2string(e_true * 1e3, precision=3, max_line_width=200))
    log("  extrinsics LMA  %s" % np.array2string(ext * 1e3, precision=3, max_line_width=200))
    log("  disparity error: before %.3f after %.3f px rms, max |d| %.2f" % (
        np.sqrt(np.mean((d0 - d_true) ** 2)), np.sqrt(np.mean((disp - d_true) ** 2)), np.abs(disp).max()))
    assert np.all(np.diff(hist) <= 1e-9 * abs(hist[0]))
    assert after < 0.2 * before


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [9, 10])
def test_lma_capture_range_with_coarse_stage(pk, orc, seed):
    """Extrinsics drawn N(0, 1e-3) put the relative port offsets about 4 px off, outside the basin
    of the zero-offset correlation loss (the LMA alone stays there).  Three coarse steps on the
    pairs' correlation-peak offsets (the reference's mismatch measurement, ID:2342-2420, solved as
    in AlignmentCorrection AC:2891-3037, 3291-3470) then seven LMA iterations bring the relative port
    offsets (what the correlations see: extrinsics and disparities as estimated, against the truth)
    below 0.25 px, and the loss at the solution is at least as low as at the true geometry: the
    remaining difference to the truth (about 0.1 px on these Bayer scenes, also with a constant
    disparity) is the bias of the loss itself, not a failure to converge.  The extrinsics alone are
    determined only up to the quad's gauge (azimuths along rXY_x and tilts along rXY_y act as a
    common disparity; a common roll barely moves relative offsets)."""
    rng = np.random.default_rng(seed)
    e_true = rng.normal(0, 1e-3, 13)
    sc = pk.synth.make_geometry_scene(512, 384, e_true, seed=seed)
    inner = (sc.tx > 1) & (sc.tx < sc.width // 8 - 2) & (sc.ty > 1) & (sc.ty < sc.height // 8 - 2)
    tiles = np.nonzero(inner)[0]
    d_true = sc.disparity_tiles.ravel()[tiles]
    d0 = (d_true + rng.normal(0, 0.05, len(tiles))).astype(np.float32)
    px, py = sc.tx[tiles] * 8 + 4.0, sc.ty[tiles] * 8 + 4.0

    def rel_err(e, d):
        a = orc.port_coords(orc.default_port_geometry(sc.width, sc.height, extrinsic=list(e)), px, py, d)
        b = orc.port_coords(orc.default_port_geometry(sc.width, sc.height, extrinsic=list(e_true)), px, py, d_true)
        rel = lambda x: np.stack([x[:, i] - x[:, j] for i, j in PAIRS], 1)
        return float(np.sqrt(np.mean((rel(a) - rel(b)) ** 2)))

    runs = {}
    for name, coarse in (("lma only", 0), ("coarse 3 + lma", 3)):
        tp = _geometry_processor(pk, sc, np.zeros(13), tiles, d0)
        ext, disp, hist = tp.lmaExtrinsics(10, coarse_iter=coarse)
        tp.close()
        runs[name] = (rel_err(ext, disp.astype(np.float64)), hist)
    tp = _geometry_processor(pk, sc, e_true, tiles, d_true.astype(np.float32))
    _, _, loss_true = tp.getGeometryGradient()
    tp.close()
    before = rel_err(np.zeros(13), d0.astype(np.float64))
    log("case: capture range, extrinsics N(0, 1e-3), seed %d: relative port offset error %.3f px at the start" % (seed, before))
    for name, (err, hist) in runs.items():
        log("  %-15s -> %.4f px; loss %.1f -> %.1f (at the truth %.1f)" % (name, err, hist[0], hist[-1], loss_true))
    assert before > 2.0
    assert runs["lma only"][0] > 1.0                       # outside the basin of the loss alone
    assert runs["coarse 3 + lma"][0] < 0.25
    assert runs["coarse 3 + lma"][1][-1] <= loss_true


def _rel_port_errors(orc, sc, tiles, e, d, e_true, d_true):
    """RMS error of the pairs' relative port offsets (what the correlations see) of extrinsics e and
    disparities d against the truth (FP64 port geometry of the oracle), px: (all of it, what the
    extrinsics alone leave once every tile's disparity is re-fitted to the truth, the worst tile)."""
    px, py = sc.tx[tiles] * 8 + 4.0, sc.ty[tiles] * 8 + 4.0

    def rel(ext, disp):
        x = orc.port_coords(orc.default_port_geometry(sc.width, sc.height, extrinsic=list(ext)), px, py,
                            np.asarray(disp, np.float64))
        return np.stack([x[:, i] - x[:, j] for i, j in PAIRS], 1).reshape(len(tiles), 12)

    truth = rel(e_true, d_true)
    r = rel(e, d) - truth
    dd = np.asarray(d, np.float64).copy()
    for _ in range(2):                    # Gauss-Newton on each tile's disparity (nearly linear)
        r0 = rel(e, dd) - truth
        j = (rel(e, dd + 1e-3) - rel(e, dd)) / 1e-3
        dd -= (j * r0).sum(1) / (j * j).sum(1)
    rx = rel(e, dd) - truth
    return (float(np.sqrt(np.mean(r ** 2))), float(np.sqrt(np.mean(rx ** 2))),
            float(np.sqrt(np.mean(r ** 2, 1)).max()))


@pytest.mark.gpu
def test_corr_peak_offsets_match_oracle(pk, orc):
    """clt_get_corr_offsets on a misaligned scene (extrinsics N(0, 1e-3) rendered, zero assumed:
    offsets of up to several px, some pairs beyond the +-7 px window) against the oracle's peak
    offsets of the same correlations: the same maxima, the centres of mass to float32 rounding."""
    rng = np.random.default_rng(4)
    sc = pk.synth.make_geometry_scene(512, 384, rng.normal(0, 1e-3, 13), seed=4)
    inner = (sc.tx > 1) & (sc.tx < sc.width // 8 - 2) & (sc.ty > 1) & (sc.ty < sc.height // 8 - 2)
    tiles = np.nonzero(inner)[0]
    d = sc.disparity_tiles.ravel()[tiles].astype(np.float32)
    tp = _geometry_processor(pk, sc, np.zeros(13), tiles, d)
    tp.execConverCorrectTiles(0)
    tp.execCorrelate2D()
    corr, _, _ = tp.getCorr2D()
    for radius in (3.9, 1.5, 7.0):
        got = tp.getCorrPeakOffsets(radius)
        want = orc.corr_peak_offsets(corr, radius)
        nan = np.isnan(want[..., 0])
        log("case: peak offsets radius %.1f, %d tasks x 6 pairs: max |offset| %.2f px, %d without a positive value, "
            "max |d| %.2g px" % (radius, len(tiles), np.nanmax(np.abs(want[..., :2])), nan.sum(),
                                 np.nanmax(np.abs(got[..., :2] - want[..., :2]))))
        assert np.array_equal(np.isnan(got[..., 0]), nan) and np.array_equal(np.isnan(got[..., 1]), nan)
        assert np.array_equal(got[..., 2], corr.reshape(len(tiles), 6, 225).max(2))
        assert np.nanmax(np.abs(got[..., :2] - want[..., :2])) < 2e-5
    with pytest.raises(RuntimeError):
        tp.getCorrPeakOffsets(0.5)
    tp.close()


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [11, 12])
def test_coarse_stage_reaches_hundredth_pixel(pk, orc, seed):
    """Extrinsics drawn N(0, 1e-3) (relative port offsets about 4.4 px off, up to 5.5 px per tile) on a
    1296x968 frame rendered through the port geometry: 10 coarse steps on the pairs' peak offsets
    leave at most 0.01 px of relative port-offset error from the extrinsics (each tile's disparity
    re-fitted to the truth), at most 0.05 px in all (the rest is the per-tile disparities, measured
    by each tile's own correlations: about 0.05 px, as precise as the 0.05 px start) and no tile
    beyond 0.5 px.  Correlation maxima at or below the strength floor 0.3, the noise peaks of pairs
    beyond the +-7 px window, are left out."""
    rng = np.random.default_rng(seed)
    e_true = rng.normal(0, 1e-3, 13)
    sc = pk.synth.make_geometry_scene(1296, 968, e_true, seed=seed)
    inner = (sc.tx > 1) & (sc.tx < sc.width // 8 - 2) & (sc.ty > 1) & (sc.ty < sc.height // 8 - 2)
    tiles = np.nonzero(inner)[0]
    d_true = sc.disparity_tiles.ravel()[tiles]
    d0 = (d_true + rng.normal(0, 0.05, len(tiles))).astype(np.float32)
    tp = _geometry_processor(pk, sc, np.zeros(13), tiles, d0)
    ext, disp, hist = tp.lmaExtrinsics(10, coarse_iter=10)
    tp.close()
    start = _rel_port_errors(orc, sc, tiles, np.zeros(13), d0, e_true, d_true)
    end = _rel_port_errors(orc, sc, tiles, ext, disp, e_true, d_true)
    log("case: coarse stage, 1296x968, extrinsics N(0, 1e-3), seed %d, %d tasks: relative port offset error "
        "%.3f -> %.4f px, from the extrinsics %.3f -> %.4f px, worst tile %.2f -> %.3f px"
        % (seed, len(tiles), start[0], end[0], start[1], end[1], start[2], end[2]))
    assert start[0] > 3.0
    assert end[1] <= 0.01 and end[0] <= 0.05 and end[2] < 0.5
