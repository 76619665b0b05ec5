"""Test helpers: run a scene through the GPU library and through the CPU oracle."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def log(msg):
    print(msg, flush=True)


def rel_rms(a, b):
    """sqrt(mean((a-b)^2)) / sqrt(mean(b^2)) over the entries where b is finite."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    m = np.isfinite(b)
    den = np.sqrt(np.mean(b[m] ** 2)) if m.any() else 0.0
    return float(np.sqrt(np.mean((a[m] - b[m]) ** 2)) / (den if den > 0 else 1.0))


def task_tiles(sc):
    """Grid tile indices of the tasks with a non-zero task word."""
    t = sc.tasks
    live = t[:, 0] != 0
    tx = t[:, 1] & 0xFFFF
    ty = (t[:, 1] >> 16) & 0xFFFF
    return (ty * (sc.width // 8) + tx)[live]


def valid_map(sc):
    v = np.zeros((sc.height // 8) * (sc.width // 8), np.uint8)
    v[task_tiles(sc)] = 1
    return v


def oracle_geom(orc, sc):
    return orc.make_geom(sc.width, sc.height, sc.num_cams, sc.num_colors, sc.red_row, sc.red_col,
                         kernels=sc.kernels is not None)


def gpu_processor(pk, sc, use_kernels=True, **kw):
    tp = pk.GPUTileProcessor(sc.width, sc.height, sc.num_cams, sc.num_colors, sc.red_row, sc.red_col,
                             kernels=sc.kernels is not None, **kw)
    if use_kernels and sc.kernels is not None:
        tp.setConvolutionKernels(sc.kernels, offsets=sc.offsets, force=True)
    tp.setBayerImages(sc.images, force=True)
    tp.setTasks(sc.tasks)
    return tp


def gpu_clt_cpu_layout(tp, sc):
    """GPU CLT of every camera [C][tiles][K][4][8][8] ([v][h] blocks, the oracle's output layout)."""
    return np.stack([tp.getCLT(c) for c in range(sc.num_cams)])


# Bars of the texture planes (relative RMS against the FP64 oracle): RGB and the rms planes as the
# reconstructed images; alpha and the port weights are exponentials of squared differences of
# nearly equal port values chosen by comparison (two best ports), where a float32 restatement of
# the reference is itself 4e-5 .. 7e-5 (alpha) and 3e-4 .. 9e-4 (weights) off.
TEX_TOL = {"rgb": 1e-4, "alpha": 2e-4, "weights": 2e-3, "rms": 1e-4}


def tex_planes(num_cams):
    return (("rgb", slice(0, 3)), ("alpha", slice(3, 4)), ("weights", slice(4, 4 + num_cams)),
            ("rms", slice(4 + num_cams, 8 + num_cams)))


def fp32_sensitive(run, clt64, moved, draws=3, seed=1, ulps=1.0):
    """Tasks whose FP64 outputs are not determined at float32 precision.

    run(clt) returns the FP64 oracle's outputs of a stage on the CLT clt [C][tiles][K][4][8][8];
    moved(out, base) returns a bool mask over the tasks of outputs moved past the bars.  The CLT is
    perturbed `draws` times by Gaussian noise of `ulps` float32 ulps of the largest coefficient of
    each (camera, tile, colour) block (2^-23 of it per ulp: about the error of a float32 transform
    chain, whose CLT is 2e-7 relative RMS off the FP64 one) and the stage rerun.  Returns (mask of
    the tasks moved by any draw, the unperturbed outputs).

    Such tasks occur where a camera's window lies mostly outside the frame: the border clamp
    leaves many bins with rounding noise only, which the phase correlation normalises to unit
    weight; and where a discrete decision of the LMA chain (a convexity test, the rounding of the
    start point, an accepted step) sits within rounding of its threshold.  No float32 computation,
    the reference's own GPU path included, can be held to the bars there."""
    import numpy as np
    rng = np.random.default_rng(seed)
    base = run(clt64)
    C, T, K = clt64.shape[:3]
    blk = clt64.reshape(C, T, K, 256)
    scale = np.abs(blk).max(-1, keepdims=True) * (ulps * 2.0 ** -23)
    bad = None
    for _ in range(draws):
        noisy = (blk + rng.standard_normal(blk.shape) * scale).reshape(clt64.shape)
        m = moved(run(noisy), base)
        del noisy
        bad = m if bad is None else (bad | m)
    return bad, base


def corr_moved(corr_tol=1e-3, disp_tol=0.01, layers=(0, 10, 1, 8, 3, 5)):
    """moved() for correlate2D outputs (corr, combo, disp): a pair correlation value moved by more
    than corr_tol, or a disparity layer (integer maximum, centre of mass, final, LMA, horizontal,
    vertical) by more than disp_tol or between NaN and finite."""
    import numpy as np

    def moved(out, base):
        n = len(base[0])
        bad = np.nan_to_num(np.abs(out[0] - base[0]).reshape(n, -1), nan=np.inf).max(1) > corr_tol
        for k in layers:
            a, b = out[2][:, k], base[2][:, k]
            fa, fb = np.isfinite(a), np.isfinite(b)
            bad |= fa != fb
            both = fa & fb
            bad[both] |= np.abs(a[both] - b[both]) > disp_tol
        return bad
    return moved


def tex_moved(bar=1e-2, planes=slice(0, 3)):
    """moved() for texture tiles [task][planes][16][16]: a pixel of `planes` (RGB) moved by more
    than `bar`."""
    import numpy as np

    def moved(out, base):
        n = len(base)
        return np.nan_to_num(np.abs(out[:, planes] - base[:, planes]).reshape(n, -1), nan=0.0).max(1) > bar
    return moved


def confirm_tex_flips(orc, g, clt64, tasks, idx, extra=None, draws=24, seed=7, bar=1e-2):
    """The subset of the tasks idx whose FP64 texture tiles move past `bar` (an RGB pixel) under
    1-ulp perturbations of the CLT, over `draws` draws of those tasks alone.  Such a tile has a
    pixel whose two-best-port choice (ID:2893-2910) is a near-tie below float32 resolution (the
    port weights of the second and third port equal to a few 1e-6): any float32 computation may
    pick either port there, and the pixel then takes the other pair's mix."""
    import numpy as np
    idx = np.asarray(idx, np.int64)
    if idx.size == 0:
        return idx
    sub = np.ascontiguousarray(tasks[idx])
    ex = None if extra is None else np.ascontiguousarray(np.asarray(extra, np.float64)[idx])
    bad, _ = fp32_sensitive(lambda c: orc.textures_frame(g, c, sub, extra_disparity=ex)[0], clt64, tex_moved(bar),
                            draws=draws, seed=seed)
    return idx[bad]


def fp32_ill_conditioned(orc, g, clt64, tasks, draws=3, seed=1, pairs=None, nthreads=None, lma=None):
    """fp32_sensitive for the oracle's correlate2D with the LMA chain: (mask, (corr, combo, disp))."""
    return fp32_sensitive(lambda c: orc.corr_frame(g, c, tasks, pairs=pairs, nthreads=nthreads, lma=lma), clt64,
                          corr_moved(), draws=draws, seed=seed)


def confirm_ill_conditioned(orc, g, clt64, tasks, idx, draws=24, seed=7, pairs=None, nthreads=None, lma=None):
    """The subset of the tasks idx that fp32_ill_conditioned marks with `draws` perturbations of
    those tasks alone (more draws than a whole-frame pass, where a few random draws can miss a
    decision that sits within rounding of its threshold)."""
    import numpy as np
    idx = np.asarray(idx, np.int64)
    if idx.size == 0:
        return idx
    sub = np.ascontiguousarray(tasks[idx])
    ill, _ = fp32_ill_conditioned(orc, g, clt64, sub, draws=draws, seed=seed, pairs=pairs, nthreads=nthreads, lma=lma)
    return idx[ill]
