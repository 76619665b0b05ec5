"""CLT kernel preparation (QuadCLT.calculateCLTKernel QC:433-668) and the kernel file formats
(flattenCLTKernels QC:670-745, extractCLTKernels QC:790-857, saveFloatKernels TQ:1076-1279).

Oracle pins (CPU), each independent of the oracle's own DCT-IIIe tables:
* a centred delta kernel becomes the identity: CC = 1 everywhere (norm_kern) or 1/16 (raw), the
  other quadrants 0, CltExtra 0;
* closed form: for a kernel whose 15x15 downsample has no taps at |i| = 7 or |j| = 7, the four
  quadrants are sums of the kernel against cos/sin at the half-integer frequencies
  w_k = (2k+1) pi / 16 divided by 16 (the 2-D DTFT of the kernel at those frequencies); SS follows
  the reference's sign pattern (ID:4717), i.e. the sine-sine sum of the x-odd part made y-odd;
  with norm_kern, CC[0][0] is exactly 1;
* the 3x3 downsample weights (1, 1/2 sides, 1/4 corners) on single odd-offset samples;
* the chosen centre, its clamps to [15, ks - 16] and the data_x/center_x items;
* the slopes of linear centre fields (exact least squares) with a linear sensor correction.
GPU parity: the device chain against the oracle on non-separable, asymmetric PSFs (FP64 on both
sides, float32 outputs: 1e-6 of the largest coefficient), and convert_direct with the prepared
kernels against the oracle's convert with its own (1e-4 relative RMS).
"""
import numpy as np
import pytest

from helpers import gpu_processor, log, oracle_geom, rel_rms

KS = 64
C0 = KS // 2


def _dtft_quadrants(k15):
    """[4][v][h] of a 15x15 kernel (rows = vertical offset -7..7) from the 2-D DTFT at the
    frequencies (2k+1) pi / 16, scaled by 1/16; SS with the reference's sign pattern."""
    m = np.arange(-7, 8)
    w = (2 * np.arange(8) + 1) * np.pi / 16
    cos_v, sin_v = np.cos(np.outer(w, m)), np.sin(np.outer(w, m))   # [freq][offset]
    cos_h, sin_h = cos_v, sin_v
    kt = np.sign(m)[:, None] * 0.5 * (k15 + k15[::-1, :])   # y-even part made y-odd
    xo = 0.5 * (kt - kt[:, ::-1])                           # ... and x-odd
    out = np.empty((4, 8, 8))
    out[0] = cos_v @ k15 @ cos_h.T
    out[1] = cos_v @ k15 @ sin_h.T
    out[2] = sin_v @ k15 @ cos_h.T
    out[3] = sin_v @ xo @ sin_h.T
    return out / 16


def _stack(kernels):
    """[K][KY][KX][ks][ks] -> the raw stack [K][KY*ks][KX*ks]."""
    K, KY, KX = kernels.shape[:3]
    return np.ascontiguousarray(kernels.transpose(0, 1, 3, 2, 4).reshape(K, KY * KS, KX * KS), np.float32)


def test_delta_kernel_is_identity(orc):
    raw = np.zeros((3, 2, 3, KS, KS))
    raw[..., C0, C0] = 1.0
    for norm, cc in ((True, 1.0), (False, 1.0 / 16)):
        kern, offs = orc.prepare_kernels(_stack(raw), 3, 2, norm_kern=norm)
        assert np.allclose(kern[..., 0, :, :], cc, rtol=0, atol=1e-7)
        assert np.abs(kern[..., 1:, :, :]).max() < 1e-7
        assert np.abs(offs).max() == 0.0


def test_quadrants_closed_form(orc):
    """Random kernels on the even lattice around the centre (so every 15x15 tap is one raw sample)
    with a dominant centre (so the chosen centre is the middle): quadrants == 2-D DTFT / 16."""
    rng = np.random.default_rng(5)
    KY, KX, K = 2, 3, 3
    raw = np.zeros((K, KY, KX, KS, KS))
    k15s = np.zeros((K, KY, KX, 15, 15))
    for idx in np.ndindex(K, KY, KX):
        k15 = np.zeros((15, 15))
        k15[1:14, 1:14] = rng.uniform(-0.3, 1.0, (13, 13)) * 0.05   # no taps at |7|
        k15[7, 7] = 8.0
        r = np.zeros((KS, KS))
        r[C0 - 14:C0 + 15:2, C0 - 14:C0 + 15:2] = k15
        ii, jj = np.mgrid[-C0:C0, -C0:C0]
        assert abs((r * jj).sum() / r.sum()) < 0.5 and abs((r * ii).sum() / r.sum()) < 0.5
        raw[idx] = r
        k15s[idx] = k15
    for norm in (False, True):
        kern, offs = orc.prepare_kernels(_stack(raw), KX, KY, norm_kern=norm)
        for k, ty, tx in np.ndindex(K, KY, KX):
            k15 = k15s[k, ty, tx]
            exp = _dtft_quadrants(k15)
            if norm:
                exp /= exp[0, 0, 0]   # normalisation: sum(k * window) = 16 <=> CC[0][0] = 1
            got = kern[ty, tx, k]
            assert np.abs(got - exp).max() < 2e-6 * np.abs(exp).max(), (norm, k, ty, tx)
            if norm:
                assert abs(got[0, 0, 0] - 1.0) < 1e-6
        # the centre stays the middle sample: data_x = data_y = 0; center = centroid / 2
        ii, jj = np.mgrid[-C0:C0, -C0:C0]
        r = raw[1, 1, 2]
        assert offs[1, 2, 1, 0] == 0 and offs[1, 2, 1, 1] == 0
        assert abs(offs[1, 2, 1, 2] - 0.5 * (r * jj).sum() / r.sum()) < 1e-6
        assert abs(offs[1, 2, 1, 3] - 0.5 * (r * ii).sum() / r.sum()) < 1e-6


def test_downsample_weights(orc):
    """A satellite at an odd offset is shared by the neighbouring 15x15 taps: 1/2 each on a side,
    1/4 each on a diagonal (ID:4649-4651)."""
    v = 0.2
    raw = np.zeros((1, 1, 2, KS, KS))
    raw[..., C0, C0] = 4.0
    raw[0, 0, 0, C0, C0 + 5] = v            # between taps j = 2 and j = 3 of row 0
    raw[0, 0, 1, C0 + 3, C0 - 3] = v        # between rows 1, 2 and columns -2, -1
    k15 = np.zeros((2, 15, 15))
    k15[:, 7, 7] = 4.0
    k15[0, 7, 9] += v / 2
    k15[0, 7, 10] += v / 2
    for i in (8, 9):
        for j in (5, 6):
            k15[1, i, j] += v / 4
    kern, _ = orc.prepare_kernels(np.ascontiguousarray(_stack(raw).reshape(1, KS, 2 * KS)), 2, 1, norm_kern=False)
    for t in range(2):
        exp = _dtft_quadrants(k15[t])
        assert np.abs(kern[0, t, 0] - exp).max() < 1e-7 * 4


def test_centre_choice_and_clamps(orc):
    """data_x = (chosen centre - ks/2) / 2 with the chosen centre clamped to [15, ks - 16]
    (ID:4646-4650; the low clamp is asymmetric); a delta moved by an even offset stays the
    identity kernel."""
    cases = [(C0 + 6, C0 - 4, 3.0, -2.0), (C0 + 22, C0, 8.0, 0.0), (5, C0, -8.5, 0.0), (C0, 3, 0.0, -8.5),
             (C0 - 16, C0 + 16, -8.0, 8.0)]
    raw = np.zeros((1, 1, len(cases), KS, KS))
    for t, (x, y, _, _) in enumerate(cases):
        raw[0, 0, t, y, x] = 1.0
    kern, offs = orc.prepare_kernels(np.ascontiguousarray(_stack(raw)), len(cases), 1, norm_kern=True)
    for t, (x, y, dx, dy) in enumerate(cases):
        e = offs[0, t, 0]
        assert (e[0], e[1]) == (dx, dy), (t, e)
        assert (e[2], e[3]) == (0.5 * (x - C0), 0.5 * (y - C0)), (t, e)
    assert np.allclose(kern[0, 0, 0, 0], 1.0, atol=1e-7) and np.allclose(kern[0, 4, 0, 0], 1.0, atol=1e-7)


def test_centre_slopes_and_sensor_offset(orc):
    """Centres linear over the kernel grid (two-point kernels with exact sub-pixel centroids) and a
    linear sensor correction: items 4..7 = d(center)/d(tile) / kernel_step, exact at the borders
    too (least squares on a plane); items 0..3 move by minus the correction (ID:4781-4791)."""
    KY, KX, step = 4, 5, 16
    ty, tx = np.mgrid[0:KY, 0:KX].astype(np.float64)
    cx = 0.02 + 0.015 * tx + 0.01 * ty        # desired center_x (sensor px) < 0.25
    cy = 0.2 - 0.02 * tx - 0.005 * ty
    raw = np.zeros((1, KY, KX, KS, KS))
    for y in range(KY):
        for x in range(KX):
            # centroid 2c (double-res px) from weights 1 at the centre and t at +2: 2t/(1+t) = 2c
            tx_ = cx[y, x] / (1 - cx[y, x])
            ty_ = cy[y, x] / (1 - cy[y, x])
            raw[0, y, x, C0:C0 + 3:2, C0:C0 + 3:2] = np.outer([1.0, ty_], [1.0, tx_])   # separable 2x2
    sxy = np.stack([0.3 + 0.05 * tx - 0.02 * ty, -0.1 + 0.03 * ty], axis=-1).astype(np.float32)
    kern, offs = orc.prepare_kernels(np.ascontiguousarray(_stack(raw)), KX, KY, kernel_step=step, sensor_xy=sxy)
    e = offs[:, :, 0].astype(np.float64)
    assert np.abs(e[..., 2] - (cx - sxy[..., 0])).max() < 1e-6
    assert np.abs(e[..., 3] - (cy - sxy[..., 1])).max() < 1e-6
    assert np.abs(e[..., 0] + sxy[..., 0]).max() < 1e-7 and np.abs(e[..., 1] + sxy[..., 1]).max() < 1e-7
    exp = [(0.015 - 0.05) / step, (0.01 + 0.02) / step, (-0.02) / step, (-0.005 - 0.03) / step]
    for i in range(4):
        assert np.abs(e[..., 4 + i] - exp[i]).max() < 1e-6, i


def _coded_kernels(KY, KX, K):
    k = np.arange(KY * KX * K * 256, dtype=np.float32).reshape(KY, KX, K, 4, 8, 8)
    o = -np.arange(KY * KX * K * 8, dtype=np.float32).reshape(KY, KX, K, 8) - 1
    return k, o


def test_flat_layout(pk):
    """flattenCLTKernels (QC:703-736): tile (tx, ty) of colour k at rows 17 ty.., columns 16 tx..;
    quadrant q at row offset 8 (q >> 1), column offset 8 (q & 1), holding the CPU-orientation
    coefficient [i][j] = the upload [v][h] block's [j][i]; extras in row 16."""
    KY, KX, K = 3, 4, 2
    k, o = _coded_kernels(KY, KX, K)
    flat = pk.kernel_files.flatten_clt_kernels(k, o)
    assert flat.shape == (K, 17 * KY, 16 * KX)
    rng = np.random.default_rng(0)
    for _ in range(200):
        ty, tx, c, q, i, j = (int(rng.integers(n)) for n in (KY, KX, K, 4, 8, 8))
        assert flat[c, 17 * ty + 8 * (q >> 1) + i, 16 * tx + 8 * (q & 1) + j] == k[ty, tx, c, q, j, i]
    assert np.array_equal(flat[:, 16::17, :].reshape(K, KY, KX, 16)[..., :8], o.transpose(2, 0, 1, 3))
    assert not flat[:, 16::17, :].reshape(K, KY, KX, 16)[..., 8:].any()
    k2, o2 = pk.kernel_files.extract_clt_kernels(flat)
    assert np.array_equal(k2, k) and np.array_equal(o2, o)


def test_float_kernel_dumps(pk, tmp_path):
    """saveFloatKernels (TQ:1076-1177) / saveFloatKernelsBigEndian (TQ:1181-1279): file names,
    sizes, byte order and the transposed (upload) / CPU orientations; read back exactly."""
    kf = pk.kernel_files
    KY, KX, K = 2, 3, 3
    ks, os_ = [], []
    for cam in range(2):
        k, o = _coded_kernels(KY, KX, K)
        ks.append(k + cam * 0.5)
        os_.append(o - cam)
    img = np.random.default_rng(1).uniform(0, 100, (2, 3, 6 * 5)).astype(np.float32)
    pxy = np.random.default_rng(2).uniform(0, 100, (7, 2, 2))
    for transpose in (True, False):
        for be in (False, True):
            prefix = str(tmp_path / ("d%d%d" % (transpose, be)))
            kf.save_float_kernels(prefix, ks, os_, img, pxy, transpose=transpose, big_endian=be)
            dt = ">f4" if be else "<f4"
            for cam in range(2):
                kp, op = kf.kernel_paths(prefix, cam, transpose)
                assert kp.endswith("_chn%d%s.kernel" % (cam, "_transposed" if transpose else ""))
                raw = np.fromfile(kp, dt).reshape(KY, KX, K, 4, 8, 8)
                assert np.array_equal(raw, ks[cam] if transpose else np.swapaxes(ks[cam], -1, -2))
                assert np.array_equal(np.fromfile(op, dt).reshape(KY, KX, K, 8), os_[cam])
                k2, o2 = kf.read_float_kernels(prefix, cam, KX, KY, K, transpose=transpose, big_endian=be)
                assert np.array_equal(k2, ks[cam]) and np.array_equal(o2, os_[cam])
                b = kf.read_bayer(prefix, cam, 6, 5, big_endian=be)
                assert np.array_equal(b.ravel(), img[cam].astype(np.float64).sum(0).astype(np.float32))
                assert np.array_equal(kf.read_ports_xy(prefix, cam, big_endian=be), pxy[:, cam].astype(np.float32))
    with pytest.raises(ValueError):
        kf.read_float_kernels(str(tmp_path / "d10"), 0, KX + 1, KY, K)


# ------------------------------------------------------------------------------------- GPU
def _prep_case(pk, orc, sc, seed, sensor, norm, cams):
    rng = np.random.default_rng(seed)
    tp = gpu_processor(pk, sc, use_kernels=False)
    want = []
    for cam in cams:
        raw = pk.synth.raw_kernel_stack(sc.num_colors, sc.kernels_vert, sc.kernels_hor, rng)
        sxy = None
        if sensor:
            sxy = np.stack([pk.synth.smooth_field(sc.kernels_vert, sc.kernels_hor, rng, -0.7, 0.7) for _ in range(2)],
                           axis=-1).astype(np.float32)
        tp.calculateCLTKernel(cam, raw, KS, sensor_xy=sxy, norm_kern=norm)
        want.append(orc.prepare_kernels(raw, sc.kernels_hor, sc.kernels_vert, sensor_xy=sxy, norm_kern=norm))
    return tp, want


@pytest.mark.gpu
@pytest.mark.parametrize("sensor,norm", [(True, True), (False, False)])
def test_gpu_prepare_kernels_parity(pk, orc, sensor, norm):
    log("case: device kernel preparation, quad 128x96 (10x8 kernel grid), sensor=%s norm=%s" % (sensor, norm))
    sc = pk.synth.make_scene(128, 96, seed=31, disparity=2.0)
    tp, want = _prep_case(pk, orc, sc, 7, sensor, norm, range(sc.num_cams))
    for cam in range(sc.num_cams):
        gk, go = tp.getConvolutionKernel(cam)
        ok, oo = want[cam]
        dk = np.abs(gk - ok).max() / np.abs(ok).max()
        do = np.abs(go - oo).max()
        log("  cam %d: kernels max|d|/max=%.3g, extras max|d|=%.3g" % (cam, dk, do))
        assert dk < 1e-6 and do < 1e-5
    # convert with the device-prepared kernels == the oracle's convert with its kernels
    tp.execConverCorrectTiles()
    g_clt = np.stack([tp.getCLT(c) for c in range(sc.num_cams)])
    tp.close()
    o_clt = orc.convert_frame(oracle_geom(orc, sc), sc.images, sc.tasks,
                              kernels=np.stack([w[0] for w in want]), offsets=np.stack([w[1] for w in want]))
    r = rel_rms(g_clt, o_clt)
    log("  convert with prepared kernels: rel RMS %.3g" % r)
    assert r < 1e-4


@pytest.mark.gpu
def test_gpu_prepare_kernels_full_grid(pk, orc):
    """2592x1936 geometry: the 164x123 kernel grid of one camera (a 991 MB raw stack, uploaded in
    four bands of kernel rows)."""
    log("case: device kernel preparation, 2592x1936 geometry, camera 2, 164x123x3 kernels")
    tp = pk.GPUTileProcessor(2592, 1936, 4, 3, 0, 1, kernels=True)
    rng = np.random.default_rng(11)
    raw = pk.synth.raw_kernel_stack(3, tp.KERNELS_VERT, tp.KERNELS_HOR, rng)
    sxy = np.stack([pk.synth.smooth_field(tp.KERNELS_VERT, tp.KERNELS_HOR, rng, -1.0, 1.0) for _ in range(2)],
                   axis=-1).astype(np.float32)
    tp.calculateCLTKernel(2, raw, KS, sensor_xy=sxy)
    gk, go = tp.getConvolutionKernel(2)
    tp.close()
    ok, oo = orc.prepare_kernels(raw, 164, 123, sensor_xy=sxy)
    dk = np.abs(gk - ok).max() / np.abs(ok).max()
    do = np.abs(go - oo).max()
    log("  kernels max|d|/max=%.3g, extras max|d|=%.3g" % (dk, do))
    assert dk < 1e-6 and do < 1e-5


@pytest.mark.gpu
def test_gpu_get_kernels_round_trip_and_errors(pk):
    sc = pk.synth.make_scene(64, 48, seed=3)
    tp = gpu_processor(pk, sc)
    for cam in range(sc.num_cams):
        k, o = tp.getConvolutionKernel(cam)
        assert np.array_equal(k, sc.kernels[cam]) and np.array_equal(o, sc.offsets[cam])
    raw = np.zeros((3, sc.kernels_vert * KS, sc.kernels_hor * KS), np.float32)
    with pytest.raises(ValueError):
        tp.calculateCLTKernel(0, raw[:, :-1], KS)
    with pytest.raises(pk.CltError):
        tp.calculateCLTKernel(9, raw, KS)
    tp.clearConvolutionKernels()
    with pytest.raises(pk.CltError):
        tp.getConvolutionKernel(0)
    tp.close()
