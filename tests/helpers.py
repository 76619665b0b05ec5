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


def robust_rel_rms(a, b, frac=1e-3, tol=1e-3):
    """For outputs of discontinuous selections (argmax/argmin of port weights in the reference's
    texture combiner): FP32 and FP64 may resolve near-ties differently on a few pixels.
    Returns (rel RMS over all but the worst `frac` of the entries, fraction of entries whose error
    exceeds `tol` times the RMS of b)."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    m = np.isfinite(b)
    a, b = a[m], b[m]
    err = np.abs(a - b)
    scale = np.sqrt(np.mean(b ** 2)) or 1.0
    keep = np.argsort(err)[: max(1, int(np.ceil(len(err) * (1 - frac))))]
    r = float(np.sqrt(np.mean(err[keep] ** 2)) / scale)
    return r, float(np.mean(err > tol * scale))
