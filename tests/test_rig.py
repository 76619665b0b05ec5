"""Two-quad rig (ImageDtt.clt_bi_quad ID:8731-9504: tileCorrs ID:7581-7740 of each quad,
correlateInterCamerasFD C2:401-448, tileInterCamCorrs ID:7370-7560, the low-texture pass
ID:9404-9480).

Oracle pins (CPU):
* quadraticMaxV2dX2Y2XY (PA:201-224 over quadraticApproximation PA:350-580): an exact quadratic
  surface gives its stationary point (numpy solve of the 2x2 gradient system); degenerate sample sets
  (collinear samples) give none;
* single2dCM (C2:1538-1586): a symmetric positive blob gives its centre exactly; negatives ignored;
* single2dPoly (C2:1509-1536): a sampled paraboloid gives its apex exactly, a minimum none;
* the inter-camera correlation of two quads whose four ports are identical copies of cameras A and B
  equals the pair correlation (A, B) of correlate2D, and its centre of mass / LMA disparity equal
  that pair's (so the inter path inherits the pair-sign pins of the quad tests);
* on synthetic rigs with a known residual disparity r, the inter disparity is r times the rig
  baseline (2 quad baselines here) within the centre-of-mass bias, with the sign of r;
* the low-texture pass equals the analysis of the cos-weighted neighbour averages;
* the notch form ignores the strip's axis row; skipped tasks give zeros and a NaN target.
GPU parity: the device rig against the oracle's analysis of the device's own correlations (1e-4,
same discrete decisions) and against the oracle end to end (no tile off by more than 0.01 px among
the tiles whose FP64 result is determined at float32 precision, helpers.fp32_sensitive).
"""
import numpy as np
import pytest

from helpers import (TEX_TOL, confirm_tex_flips, fp32_sensitive, gpu_clt_cpu_layout, gpu_processor, log, oracle_geom,
                     rel_rms, tex_moved, tex_planes)

RIG_OFF = np.array([2.0, 0.0])   # aux quad centre, in main-quad disparity units
AUX_SCALE = 0.75                 # aux baseline / main baseline


def rig_rxy(pk):
    q = pk.geometry.QUAD_RXY
    return np.concatenate([q, RIG_OFF + AUX_SCALE * q])


def rig_scene(pk, w, h, seed, **kw):
    return pk.synth.make_scene(w, h, num_cams=8, rxy=rig_rxy(pk), seed=seed, **kw)


# ------------------------------------------------------------------------------------- oracle pins
def test_quadratic_max_closed_form(orc):
    rng = np.random.default_rng(0)
    x, y = rng.uniform(-3, 3, 50), rng.uniform(-2, 2, 50)
    A, B, C, D, E, F = -0.5, -0.3, 0.1, 0.4, -0.2, 2.0
    z = A * x * x + B * y * y + C * x * y + D * x + E * y + F
    w = rng.uniform(0.1, 2.0, 50)
    q = orc.quadratic_max(x, y, z, w)
    xy = np.linalg.solve([[2 * A, C], [C, 2 * B]], [-D, -E])
    v = A * xy[0] ** 2 + B * xy[1] ** 2 + C * xy[0] * xy[1] + D * xy[0] + E * xy[1] + F
    assert np.allclose(q, [xy[0], xy[1], v, A, B], atol=1e-10)
    assert orc.quadratic_max(x, np.zeros(50), z, w) is None              # collinear samples
    assert orc.quadratic_max(x, 0.5 * x + 1, z, w) is None
    w0 = w.copy()
    w0[:] = 0
    assert orc.quadratic_max(x, y, z, w0) is None                        # no weight


def test_single2d_cm_blob(orc):
    c = -0.05 * np.ones((15, 15))
    for (yy, xx), v in (((-1, 1), 1.0), ((-1, 0), 0.4), ((-1, 2), 0.4), ((0, 1), 0.3), ((-2, 1), 0.3)):
        c[7 + yy, 7 + xx] = v
    out = orc.single2d_cm(c, 0.0, 3.9)
    assert np.allclose(out, [-1.0, 1.0, 1.0], atol=1e-12)


@pytest.mark.parametrize("x0,y0", [(0.3, -0.4), (-1.2, 0.7), (0.0, 0.0)])
def test_single2d_poly_paraboloid(orc, x0, y0):
    yy, xx = np.mgrid[-7:8, -7:8].astype(np.float64)
    c = 2.0 - 0.11 * (xx - x0) ** 2 - 0.07 * (yy - y0) ** 2 - 0.02 * (xx - x0) * (yy - y0)
    out = orc.single2d_poly(c, 0.0)
    assert np.allclose(out[:2], [-x0, -y0], atol=1e-9) and abs(out[2] - 2.0) < 1e-9
    out = orc.single2d_poly(-c, 0.0)                                      # a minimum: no maximum
    assert np.isnan(out[0]) and np.isnan(out[1]) and out[2] == 0.0


def _copies_clt(orc, pk, seed):
    """An 8-camera CLT whose main ports are all camera A of a horizontal pair and whose aux
    ports are all camera B, plus the pair's own correlate2D."""
    sc = pk.synth.make_scene(64, 48, num_cams=2, rxy=np.array([[-0.5, 0.0], [0.5, 0.0]]), seed=seed,
                             disparity=("field", 0.5, 2.5), target_disparity=0.0)
    g2 = oracle_geom(orc, sc)
    clt2 = orc.convert_frame(g2, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets)
    clt8 = np.concatenate([np.repeat(clt2[:1], 4, axis=0), np.repeat(clt2[1:], 4, axis=0)])
    tasks8 = np.zeros((len(sc.tasks), 2 + 16), np.int32)
    tasks8[:, :2] = sc.tasks[:, :2]
    return sc, g2, clt2, clt8, tasks8


def test_inter_is_the_pair_correlation_of_the_means(orc, pk):
    sc, g2, clt2, clt8, tasks8 = _copies_clt(orc, pk, 3)
    g8 = orc.make_geom(64, 48, 8, 3, kernels=False)
    lp = orc.default_corr_lma_params(fo_correct=0)
    r = orc.rig_frame(g8, clt8, tasks8, rig=orc.default_rig_params(use_poly=0, no_int_x0=0), lma=lp)
    corr, _, disp = orc.corr_frame(g2, clt2, sc.tasks, pairs=[(0, 1, 0, 1)], lma=lp)
    assert np.abs(r["inter"] - corr[:, 0]).max() < 1e-12 * np.abs(corr).max()
    cm = disp[:, orc.DISP_CM_RAW]
    fin = np.isfinite(cm)
    assert fin.mean() > 0.9
    assert np.abs(r["bimap"][fin, orc.BI_DISP_CROSS] - cm[fin]).max() < 1e-9
    r2 = orc.rig_frame(g8, clt8, tasks8, rig=orc.default_rig_params(no_int_x0=0), lma=lp)
    both = np.isfinite(disp[:, orc.DISP_POLY]) & np.isfinite(r2["bimap"][:, orc.BI_DISP_CROSS_DX])
    assert both.sum() > 0.5 * len(cm)
    assert np.abs(r2["bimap"][both, orc.BI_DISP_CROSS] - disp[both, orc.DISP_POLY]).max() < 1e-9


def test_tile_corrs_is_the_lma_chain_without_foreground(orc, pk):
    sc = rig_scene(pk, 64, 48, 4, disparity=("field", 1.0, 3.0), target_disparity=1.5)
    g8 = orc.make_geom(64, 48, 8, 3)
    clt = orc.convert_frame(g8, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets)
    r = orc.rig_frame(g8, clt, sc.tasks)
    g4 = orc.make_geom(64, 48, 4, 3)
    for q, (bd, bs) in enumerate(((orc.BI_DISP_FULL, orc.BI_STR_FULL), (orc.BI_ADISP_FULL, orc.BI_ASTR_FULL))):
        tasks4 = np.concatenate([sc.tasks[:, :2], sc.tasks[:, 2 + 8 * q:10 + 8 * q]], axis=1)
        _, _, d = orc.corr_frame(g4, clt[4 * q:4 * q + 4], tasks4, lma=orc.default_corr_lma_params(fo_correct=0))
        assert np.array_equal(np.isnan(d[:, orc.DISP_CM]), np.isnan(r["bimap"][:, bd]))
        f = np.isfinite(d[:, orc.DISP_CM])
        assert np.abs(d[f, orc.DISP_CM] - r["bimap"][f, bd]).max() < 1e-12
        assert np.abs(d[:, orc.DISP_STRENGTH] - r["bimap"][:, bs]).max() < 1e-12
    # one-direction fits exist exactly where the fit over all pairs succeeded
    fitted = np.isfinite(r["bimap"][:, orc.BI_DISP_HOR]) | np.isfinite(r["bimap"][:, orc.BI_DISP_VERT])
    assert fitted.mean() > 0.5
    assert (r["bimap"][~np.isfinite(r["bimap"][:, orc.BI_DISP_HOR]), orc.BI_STR_HOR] == 0).all()
    b = r["bimap"]
    assert np.allclose(b[:, orc.BI_STR_ALL], np.cbrt(b[:, orc.BI_STR_CROSS] * b[:, orc.BI_STR_FULL] *
                                                       b[:, orc.BI_ASTR_FULL]), equal_nan=True)


@pytest.mark.parametrize("residual", [0.5, -0.5])
def test_rig_residual_disparity(orc, pk, residual):
    sc = rig_scene(pk, 128, 96, 5, disparity=3.0, target_disparity=3.0 - residual)
    g8 = orc.make_geom(128, 96, 8, 3)
    clt = orc.convert_frame(g8, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets)
    b = orc.rig_frame(g8, clt, sc.tasks)["bimap"]
    cross = np.nanmedian(b[:, orc.BI_DISP_CROSS])
    main = np.nanmedian(b[:, orc.BI_DISP_FULL])
    log("  residual %.2f: main %.3f, aux %.3f, cross %.3f (expected %.2f)" %
        (residual, main, np.nanmedian(b[:, orc.BI_ADISP_FULL]), cross, residual * RIG_OFF[0]))
    assert np.sign(main) == np.sign(residual) and np.sign(cross) == np.sign(residual)
    assert abs(cross - residual * RIG_OFF[0]) < 0.3 * abs(residual * RIG_OFF[0])
    assert abs(np.nanmedian(b[:, orc.BI_DISP_CROSS_DX]) - residual * RIG_OFF[0]) < 0.3 * abs(residual * RIG_OFF[0])


def test_low_texture_pass(orc, pk):
    sc = rig_scene(pk, 64, 48, 6, disparity=("field", 1.0, 3.0), target_disparity=2.0, task_fraction=0.8)
    g8 = orc.make_geom(64, 48, 8, 3)
    clt = orc.convert_frame(g8, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets)
    r = orc.rig_frame(g8, clt, sc.tasks)
    lt = orc.rig_analysis(g8, sc.tasks, r["corr_main"], r["disp_main"], r["corr_aux"], r["disp_aux"], r["inter"],
                          rig=orc.default_rig_params(lt_rad=1))
    # the cos-weighted average of the neighbours' inter correlations (ID:9404-9454)
    tx, ty = sc.tasks[:, 1] & 0xFFFF, sc.tasks[:, 1] >> 16
    where = {(x, y): i for i, (x, y) in enumerate(zip(tx, ty))}
    avg = np.zeros_like(r["inter"])
    for i, (x, y) in enumerate(zip(tx, ty)):
        acc, sw = 0.0, 0.0
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                j = where.get((x + dx, y + dy))
                if j is None:
                    continue
                wgt = np.cos(np.pi * abs(dy) / 3) * np.cos(np.pi * abs(dx) / 3)
                acc = acc + wgt * r["inter"][j]
                sw += wgt
        avg[i] = acc / sw
    ref = orc.rig_analysis(g8, sc.tasks, r["corr_main"], r["disp_main"], r["corr_aux"], r["disp_aux"], avg,
                           rig=orc.default_rig_params(no_int_x0=1))
    for k in (orc.BI_DISP_CROSS, orc.BI_STR_CROSS, orc.BI_DISP_CROSS_DX, orc.BI_DISP_CROSS_DY, orc.BI_STR_ALL):
        assert np.allclose(lt[:, k], ref[:, k], atol=1e-9, equal_nan=True), k


def test_notch_ignores_the_axis_row_and_skipped_tasks(orc, pk):
    sc = rig_scene(pk, 32, 24, 7, disparity=2.0)
    g8 = orc.make_geom(32, 24, 8, 3)
    T = len(sc.tasks)
    inter = np.zeros((T, 15, 15))
    inter[:, 7, :] = np.exp(-0.5 * ((np.arange(15) - 9) / 1.0) ** 2)   # the axis row only, maximum at x = +2
    cm = np.zeros((T, 6, 15, 15))
    d = np.full((T, orc.DISP_LAYERS), np.nan)
    d[:, orc.DISP_STR_RAW] = 0.0
    tasks = sc.tasks.copy()
    tasks[0, 0] = 0
    b = orc.rig_analysis(g8, tasks, cm, d, cm, d, inter, rig=orc.default_rig_params(notch_mode=1))
    assert np.allclose(b[1:, orc.BI_DISP_CROSS], -2.0) and np.allclose(b[1:, orc.BI_STR_CROSS], 1.0)
    assert np.isnan(b[1:, orc.BI_DISP_CROSS_DX]).all()
    assert np.isnan(b[0, orc.BI_TARGET]) and (b[0, :orc.BI_TARGET] == 0).all()
    b = orc.rig_analysis(g8, tasks, cm, d, cm, d, inter, rig=orc.default_rig_params(no_int_x0=0))
    assert np.allclose(b[1:, orc.BI_DISP_CROSS], -2.0, atol=1e-9)   # symmetric row: the centre of mass is on +2


# ------------------------------------------------------------------------------------- GPU
LAYERS = {"disp_full": 0, "disp_hor": 1, "disp_vert": 2, "adisp_full": 5, "cross": 10, "cross_dx": 11, "cross_dy": 12,
          "str_full": 13, "astr_full": 18, "str_cross": 23, "str_all": 24}


@pytest.mark.gpu
@pytest.mark.parametrize("name,rig,fz", [("defaults", {}, 0.0),
                                         ("lt_rad 1, CM dx/dy", dict(lt_rad=1, use_xy_poly=0), 0.0),
                                         ("notch", dict(notch_mode=1), 0.0),
                                         ("fat zero 0.05, no_int_x0 off", dict(no_int_x0=0), 0.05)])
def test_gpu_rig_parity(pk, orc, name, rig, fz):
    log("case: two-quad rig 128x96, %s" % name)
    sc = rig_scene(pk, 128, 96, 8, disparity=("field", 1.0, 4.0), target_disparity=2.0)
    tp = gpu_processor(pk, sc)
    tp.setCorrSigma(0.8)
    if fz:
        tp.setCorrParams(fat_zero=fz)
    if rig:
        tp.setRigParams(**rig)
    tp.execConverCorrectTiles()
    tp.execCorrelate2DRig(np.full(len(sc.tasks), 2.0, np.float32))
    g = tp.getCorr2DRig()
    tp.close()
    rp = orc.default_rig_params(**rig)
    prm = orc.default_corr_params(fat_zero=fz)
    g8 = oracle_geom(orc, sc)
    same = orc.rig_analysis(g8, sc.tasks, g["corr_main"], g["disp_main"], g["corr_aux"], g["disp_aux"], g["inter"],
                            target=np.full(len(sc.tasks), 2.0, np.float32), params=prm, rig=rp)
    gb = g["bimap"].astype(np.float64)
    for nm, k in LAYERS.items():
        fo, fg = np.isfinite(same[:, k]), np.isfinite(gb[:, k])
        both = fo & fg
        d = float(np.abs(gb[both, k] - same[both, k]).max()) if both.any() else 0.0
        log("  same input %-10s nan mismatch %d, max|d| %.3g" % (nm, int((fo != fg).sum()), d))
        assert (fo != fg).sum() == 0 and d < 1e-4, nm
    clt = orc.convert_frame(g8, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets)
    e_layers = [LAYERS[nm] for nm in ("disp_full", "adisp_full", "cross", "cross_dx")]

    def moved(out, base):   # an end-to-end layer moved past 0.01 px or between NaN and finite
        bad = np.zeros(len(base), bool)
        for k in e_layers:
            fa, fb = np.isfinite(out[:, k]), np.isfinite(base[:, k])
            bad |= fa != fb
            both = fa & fb
            bad[both] |= np.abs(out[both, k] - base[both, k]) > 0.01
        return bad

    ill, e = fp32_sensitive(lambda c: orc.rig_frame(g8, c, sc.tasks, params=prm, rig=rp)["bimap"], clt, moved)
    ok = ~ill
    log("  end to end: %d of %d tasks not determined at float32 precision" % (ill.sum(), len(ill)))
    assert ill.sum() <= 0.02 * len(ill)
    for nm in ("disp_full", "adisp_full", "cross", "cross_dx"):
        k = LAYERS[nm]
        if nm == "cross_dx" and rig.get("notch_mode"):   # the notch form runs no LMA: no dx, dy
            assert np.isnan(e[:, k]).all() and np.isnan(gb[:, k]).all()
            continue
        both = np.isfinite(e[:, k]) & np.isfinite(gb[:, k])
        nan_mis = int((np.isfinite(e[ok, k]) != np.isfinite(gb[ok, k])).sum())
        sel = both & ok
        off = int((np.abs(gb[sel, k] - e[sel, k]) > 0.01).sum())
        med = float(np.median(np.abs(gb[both, k] - e[both, k]))) if both.any() else 0.0
        log("  end to end %-10s %d both, %d off by > 0.01 px, %d NaN mismatches, median %.2g" % (
            nm, both.sum(), off, nan_mis, med))
        assert both.sum() > 0.5 * len(sc.tasks) and off == 0 and nan_mis == 0, nm


@pytest.mark.gpu
def test_gpu_rig_rejects_other_camera_counts(pk):
    sc = pk.synth.make_scene(64, 48, seed=2)
    tp = gpu_processor(pk, sc)
    tp.execConverCorrectTiles()
    with pytest.raises(pk.CltError):
        tp.execCorrelate2DRig()
    tp.close()


@pytest.mark.gpu
def test_gpu_rig_textures(pk, orc):
    """generateTextureTiles of each quad (ID:9360-9398) against the oracle's textures of the same
    cameras, on the device CLT, with the quad's device disparity as the extra disparity.  A tile with
    an RGB pixel off by more than 1e-2 must be a confirmed float32 near-tie of the two-best-port
    choice (confirm_tex_flips); at most 2 % of the tiles may be, and the bars hold on the rest."""
    log("case: two-quad rig 128x96 textures of both quads")
    sc = rig_scene(pk, 128, 96, 9, disparity=("field", 1.0, 4.0), target_disparity=2.0)
    tp = gpu_processor(pk, sc)
    tp.setCorrSigma(0.8)
    tp.execConverCorrectTiles()
    tp.execCorrelate2DRig()
    b = tp.getCorr2DRig()["bimap"]
    tp.execTexturesRig()
    (tm, ta), (sm, sa) = tp.getTexturesRig()
    clt = gpu_clt_cpu_layout(tp, sc).astype(np.float64)
    tp.close()
    g4 = orc.make_geom(128, 96, 4, 3)
    for q, (tiles, stats, k) in enumerate(((tm, sm, orc.BI_DISP_FULL), (ta, sa, orc.BI_ADISP_FULL))):
        extra = np.nan_to_num(b[:, k].astype(np.float64))
        tasks4 = np.ascontiguousarray(np.concatenate([sc.tasks[:, :2], sc.tasks[:, 2 + 8 * q:10 + 8 * q]], axis=1))
        o_tiles, o_stats = orc.textures_frame(g4, clt[4 * q:4 * q + 4], tasks4, extra_disparity=extra)
        over = np.nonzero(tex_moved()(tiles, o_tiles))[0]
        flips = confirm_tex_flips(orc, g4, clt[4 * q:4 * q + 4], tasks4, over, extra=extra)
        keep = np.ones(len(tiles), bool)
        keep[over] = False
        res = {}
        for name, sl in tex_planes(4):
            res[name] = rel_rms(tiles[keep, sl], o_tiles[keep, sl])
        res["stats"] = rel_rms(stats[keep], o_stats[keep])
        log("  quad %d: " % q + ", ".join("%s=%.3g" % kv for kv in res.items()) +
            "; tiles past 1e-2: %s, confirmed near-ties: %s" % (over.tolist(), flips.tolist()))
        assert np.array_equal(np.sort(flips), over) and len(over) <= 0.02 * len(tiles)
        for name, tol in TEX_TOL.items():
            assert res[name] < tol, name
        assert res["stats"] < 1e-4



# ------------------------------------------------------------------------------------- rig geometry
def _geoms(orc, **aux_kw):
    main = orc.default_port_geometry(640, 480)
    aux = orc.default_port_geometry(640, 480, **aux_kw)
    return main, aux


def test_rig_ports_identity(orc):
    """An auxiliary camera equal to the main one, no baseline, no rig rotation: the same ports."""
    main, aux = _geoms(orc)
    main.distortion[1] = aux.distortion[1] = 0.02
    main.extrinsic[3] = aux.extrinsic[3] = 1e-3
    rig = orc.RigOffset(baseline=0.0, main_disparity_radius=150.0, aux_disparity_radius=150.0)
    px, py, d = np.array([100.0, 333.0]), np.array([50.0, 401.0]), np.array([0.0, 7.5])
    xy = orc.port_coords_rig(main, aux, rig, px, py, d)
    assert np.array_equal(xy[:, :4], xy[:, 4:])
    single = orc.port_coords(main, px, py, d)
    assert np.allclose(xy[:, :4].reshape(2, 8), np.asarray(single).reshape(2, 8), rtol=0, atol=1e-12)


def test_rig_ports_closed_form(orc):
    """No distortion, no CorrVector rotations: the auxiliary port of sensor i is
    R_roll(zoom) . (p - d R_a/R_m rXY_aux[i]) + pXY0_aux with rXY_aux the rolled quad offsets plus the
    baseline vector over R_a (GC:728-745, 770-803, 2159-2230); an aux azimuth alpha projects
    (x, y, f) through the rotation about the vertical axis."""
    B, th, rho, zm, Rm, Ra = 1256.0, 0.3, 0.02, 0.01, 150.0, 120.0
    main, aux = _geoms(orc, focal_length=5.0, pxy0=[(300.0, 250.0)] * 4)
    rig = orc.RigOffset(baseline=B, aux_angle=th, aux_roll=rho, aux_zoom=zm, main_disparity_radius=Rm,
                        aux_disparity_radius=Ra)
    px, py, d = np.array([400.0, 123.0]), np.array([200.0, 377.0]), np.array([2.0, 0.5])
    xy = orc.port_coords_rig(main, aux, rig, px, py, d)
    q = np.array(orc.QUAD_PORT_OFFSETS)
    c, s = np.cos(rho), np.sin(rho)
    rxy = np.stack([B * np.cos(th) / Ra + c * q[:, 0] + s * q[:, 1], B * np.sin(th) / Ra - s * q[:, 0] + c * q[:, 1]], 1)
    for t in range(2):
        v = np.array([px[t] - 320.0, py[t] - 240.0])[None, :] - d[t] * Ra / Rm * rxy
        cz, sz = c * (1 + zm), s * (1 + zm)
        exp = np.stack([cz * v[:, 0] + sz * v[:, 1], -sz * v[:, 0] + cz * v[:, 1]], 1) + 300.0 * np.array([1.0, 250 / 300])
        assert np.allclose(xy[t, 4:], exp, rtol=0, atol=1e-9)
    al = 0.01
    rig = orc.RigOffset(baseline=B, aux_azimuth=al, main_disparity_radius=Rm, aux_disparity_radius=Ra)
    xy = orc.port_coords_rig(main, aux, rig, px, py, d)
    fl = 5.0 / (0.001 * 2.2)
    rxy = np.stack([B / Ra + q[:, 0], q[:, 1]], 1)
    for t in range(2):
        v = np.array([px[t] - 320.0, py[t] - 240.0])[None, :] - d[t] * Ra / Rm * rxy
        rz = np.sin(al) * v[:, 0] + np.cos(al) * fl
        ex = (np.cos(al) * v[:, 0] - np.sin(al) * fl) * fl / rz + 300.0
        ey = v[:, 1] * fl / rz + 250.0
        assert np.allclose(xy[t, 4:, 0], ex, atol=1e-9) and np.allclose(xy[t, 4:, 1], ey, atol=1e-9)


@pytest.mark.gpu
def test_gpu_rig_port_centres(pk, orc):
    log("case: rig port centres on the device, 640x480, distortion, extrinsics, rig offsets")
    main, aux = _geoms(orc, focal_length=4.6, pxy0=[(322.0, 238.5), (318.2, 241.0), (320.5, 239.9), (321.1, 240.2)])
    for g in (main, aux):
        g.distortion[0], g.distortion[1], g.distortion[2] = 0.01, -0.03, 0.015
    main.extrinsic[0], main.extrinsic[4], main.extrinsic[7], main.extrinsic[11] = 2e-4, -3e-4, 1e-3, 5e-4
    aux.extrinsic[1], aux.extrinsic[3], aux.extrinsic[8], aux.extrinsic[10] = -1e-4, 2e-4, -5e-4, 3e-4
    kw = dict(baseline=1256.0, aux_angle=0.05, aux_azimuth=2e-3, aux_tilt=-1e-3, aux_roll=4e-3, aux_zoom=1e-3,
              main_disparity_radius=150.0, aux_disparity_radius=120.0)
    rig = orc.RigOffset(**kw)
    tp = pk.GPUTileProcessor(640, 480, 8, 3, kernels=False)
    gm = pk.GPUTileProcessor.defaultPortGeometry(640, 480)
    ga = pk.GPUTileProcessor.defaultPortGeometry(640, 480)
    for dst, src in ((gm, main), (ga, aux)):
        for f, _ in src._fields_:
            v = getattr(src, f)
            if hasattr(v, "__len__"):
                for i in range(len(v)):
                    if hasattr(v[i], "__len__"):
                        for j in range(len(v[i])):
                            getattr(dst, f)[i][j] = v[i][j]
                    else:
                        getattr(dst, f)[i] = v[i]
            else:
                setattr(dst, f, v)
    tp.setPortGeometryRig(gm, ga, pk.GPUTileProcessor.rigOffset(**kw))
    rng = np.random.default_rng(3)
    n = 600
    tx, ty = rng.integers(0, 80, n), rng.integers(0, 60, n)
    d = rng.uniform(0, 20, n).astype(np.float32)
    words = np.full(n, 0xff, np.int32)
    words[::7] = 0
    tp.getTileSubcamOffsets(words, tx, ty, d)
    got = tp.getTasks()[:, 2:].view(np.float32).reshape(n, 8, 2)
    tp.close()
    exp = orc.port_coords_rig(main, aux, rig, tx * 8 + 4.0, ty * 8 + 4.0, d.astype(np.float64)).astype(np.float32)
    live = words != 0
    log("  max |d| %.3g px, identical %d of %d" % (np.abs(got[live] - exp[live]).max(),
                                                    int((got[live] == exp[live]).all(axis=(1, 2)).sum()), live.sum()))
    assert np.array_equal(got[live], exp[live])
    assert (got[~live] == 0).all()
