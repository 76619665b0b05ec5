"""Benchmark of the CLT tile processor on one MI355X.

Workload (one "step" = one frame through the plain per-frame GPU path): 4 x 2592x1936 Bayer
(RGGB) float32 mosaics, 78,408 tiles, dense task list, per-tile disparity a smooth random field
0..20 px, random per-tile aberration kernels; convert_direct and correlate2D (6 pairs + combined
strip + disparity, per-tile LMA chain) in one kernel (clt_exec_convert_corr2d: the CLT passes from
one to the other in LDS) -> textures (RGBA) and imclt_rbg (with LPF) in one pass over the converted
tiles (clt_exec_textures_imclt).  The LMA chain (on its own stream) overlaps the textures/imclt
pass and the next frame's conversion.  Inputs are resident in HBM before the timed region; outputs
stay on the device.

Prints one JSON line: tiles/s, HBM roofline of the dominant kernel (HIP events on the library's
stream), the CPU baseline (the C restatement of the reference CPU path in FP64 and FP32 on the
host cores) and the error of each output against that FP64 run.
"""
import argparse
import csv
import glob
import importlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CLT tiles/s per frame (convert+corr+imclt), % HBM roofline; x vs Java CPU"
HBM_PEAK_GBS = 8000.0
W, H, C, K = 2592, 1936, 4, 3
RED = (0, 0)   # Bayer RGGB: red at even rows and even columns


def algorithmic_bytes(sc, planes, n_pairs, strip_h):
    """Compulsory bytes at each kernel's interface (f32), per launch."""
    T = len(sc.tasks)
    tw = 4 * (2 + 2 * C)
    clt = C * T * K * 1024
    # kernel tiles actually used: round(centre/16)+1, clamped (ImageDtt.java:4906-4911)
    xy = sc.tasks[:, 2:].view(np.float32).reshape(T, C, 2).astype(np.float64)
    ktx = np.clip(np.floor(xy[..., 0] / 16 + 0.5).astype(int) + 1, 0, sc.kernels_hor - 1)
    kty = np.clip(np.floor(xy[..., 1] / 16 + 0.5).astype(int) + 1, 0, sc.kernels_vert - 1)
    used = np.unique((np.arange(C)[None, :] * sc.kernels_vert + kty) * sc.kernels_hor + ktx).size
    return {
        "convert": C * W * H * 4 + used * K * (256 + 8) * 4 + T * tw + clt,
        "corr2d": clt + T * tw + T * n_pairs * 225 * 4 + T * strip_h * 15 * 4 + T * 12 * 4,
        # convert and correlate2D in one kernel: the CLT written once (for the textures), not read back
        "convert_corr2d": C * W * H * 4 + used * K * (256 + 8) * 4 + T * tw + clt + T * n_pairs * 225 * 4
                          + T * strip_h * 15 * 4 + T * 12 * 4,
        "textures": clt + T * tw + T * planes * 1024 + T * 4 * C * 4,
        "imclt": clt + C * K * (W + 8) * (H + 8) * 4 + (W // 8) * (H // 8) * 4,
        # textures and imclt_rbg in one pass: the CLT once, both outputs, the tile map
        "tex_imclt": clt + T * tw + T * planes * 1024 + T * 4 * C * 4 + C * K * (W + 8) * (H + 8) * 4
                     + (W // 8) * (H // 8) * 4,
    }


def host_cores():
    """Host threads the CPU baseline uses: the cores this process may run on, at most 16 (one
    GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def java_probe():
    """The reference CPU path is Java (ImageDtt) and needs a JVM plus the ij 1.52e and Jama 1.0.3
    jars, which the reference does not ship; the run records what it found."""
    java = shutil.which("java")
    return {"java": java, "reference_jars": None,
            "ran": False, "note": "no JVM" if java is None else "JVM present, ij/Jama jars absent"}


def oracle_frame(orc, sc, precision, cores):
    """The FP64 (or FP32) oracle of one frame: convert, correlate2D with the LMA chain, textures
    (RGBA), imclt_rbg with the LPF, as the GPU step computes them."""
    g = orc.make_geom(W, H, red_row=RED[0], red_col=RED[1])
    valid = np.zeros((H // 8) * (W // 8), np.uint8)
    valid[((sc.tasks[:, 1] >> 16) * (W // 8) + (sc.tasks[:, 1] & 0xFFFF))] = 1
    clt = orc.convert_frame(g, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets, nthreads=cores,
                            precision=precision)
    corr, combo, disp = orc.corr_frame(g, clt, sc.tasks, nthreads=cores, precision=precision)
    tex, _ = orc.textures_frame(g, clt, sc.tasks, tex_params=orc.default_tex_params(keep_weights=0), nthreads=cores,
                                precision=precision)
    img = orc.imclt_frame(g, clt, valid, lpf64=orc.lpf(0.8), nthreads=cores, precision=precision)
    return {"clt": clt, "corr": corr, "combo": combo, "disp": disp, "tex": tex, "img": img}


def cpu_baseline(sc, passes=5):
    """The oracle (C restatement of the reference CPU path, one worker per host core pulling tiles
    from an atomic counter as ImageDtt.java:5847-5872) on the same frame: FP64 and FP32 builds,
    one warm-up then `passes` timed passes each, the median reported.  Returns the baseline and
    the FP64 and FP32 outputs of the last passes."""
    import oracle as orc
    cores = host_cores()
    legs, outs = {}, {}
    for prec in (64, 32):
        oracle_frame(orc, sc, prec, cores)   # warm-up
        times = []
        for _ in range(passes):
            t0 = time.perf_counter()
            out = oracle_frame(orc, sc, prec, cores)
            times.append(time.perf_counter() - t0)
        outs[prec] = out
        med = float(np.median(times))
        legs["fp%d" % prec] = {"tiles_per_s": round(len(sc.tasks) / med, 1), "median_s": round(med, 4),
                               "passes_s": [round(t, 4) for t in times]}
    return ({"value": legs["fp64"]["tiles_per_s"], "unit": "tiles/s", "cores": cores, "kind": "port",
             "cpu_model": cpu_model(), "legs": legs, "reference_probe": java_probe(),
             "sample": "the same 78408-tile frame, convert+correlate2D(LMA)+textures+imclt; per build 1 warm-up "
                       "and the median of %d passes; value = the FP64 build" % passes}, outs)


def _rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    m = np.isfinite(b)
    den = np.sqrt(np.mean(b[m] ** 2)) or 1.0
    return float(np.sqrt(np.mean((a[m] - b[m]) ** 2)) / den)


def output_errors(gpu, ref):
    """Error of each output against a reference run of the oracle: relative RMS of the CLT, pair
    correlations, combined strips, texture tiles and images; max |difference| (px) and the count
    of tiles past 0.01 px for the disparities (integer maximum, centre of mass before the LMA, the
    final disparity after the LMA chain, the LMA disparity)."""
    e = {}
    for k in ("clt", "corr", "combo", "tex", "img"):
        e[k + "_rel_rms"] = float("%.3g" % _rel_rms(np.asarray(gpu[k]).reshape(ref[k].shape), ref[k]))
    for name, k in (("int", 0), ("cm", 10), ("final", 1), ("lma", 8)):
        g, o = gpu["disp"][:, k].astype(np.float64), ref["disp"][:, k]
        both = np.isfinite(g) & np.isfinite(o)
        d = np.abs(g[both] - o[both])
        e["disp_" + name] = {"max_px": float("%.3g" % (d.max() if d.size else 0.0)),
                             "over_0.01": int((d > 0.01).sum()),
                             "nan_mismatch": int((np.isfinite(g) != np.isfinite(o)).sum())}
    return e


def gpu_outputs(tp, sc):
    """The outputs of the last step, fetched from the device (outside the timed region)."""
    clt = np.stack([tp.getCLT(c) for c in range(C)])
    corr, combo, disp = tp.getCorr2D()
    tex, _ = tp.getTextures()
    img = np.stack([tp.getRBG(c) for c in range(C)]).reshape(C, K, H + 8, W + 8)
    return {"clt": clt, "corr": corr, "combo": combo, "disp": disp, "tex": tex, "img": img}


def copy_calibration(mib):
    """One hipMemcpy device-to-device of a known size: its FETCH_SIZE / WRITE_SIZE counters give the
    scale between the counters and real HBM bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print("bench: copied %d MiB device to device" % mib, file=sys.stderr, flush=True)


# kernel name patterns (demangled names of the counter runs)
KERNEL_NAMES = {"convert": r"k_convert\(", "corr2d": r"k_corr2d(_quad)?<", "textures": r"k_textures<",
                "imclt": r"k_imclt\(", "tex_imclt": r"k_tex_imclt<", "convert_corr2d": r"k_convert_corr<"}
CHILD_FRAMES = 2   # frames of the profiled child run
COPY_MIB = 256
STEP_KERNELS = ("convert_corr2d", "tex_imclt")   # the calls of one step (their event timings)


def pmc_traffic(seed):
    """HBM bytes per launch of each kernel from two rocprofv3 counter runs (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch) of a short child run of this script.  Each counter is scaled by
    what it reads for a device-to-device copy of COPY_MIB MiB in the same child run.  Returns
    ({kernel: bytes}, {"fetch": factor, "write": factor}) or (None, reason).  A call of the one-pass
    textures/imclt is two dispatches (even, then odd tile columns): bytes are per call."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return None, "rocprofv3 not found"
    vals, factors = {}, {}
    out_dir = tempfile.mkdtemp(prefix="clt_pmc_")
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            name = counter.lower()
            cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", out_dir, "-o", name, "--",
                   sys.executable, os.path.abspath(__file__), "--pmc-child", "--seed", str(seed)]
            print("bench: counter run %s ..." % counter, file=sys.stderr, flush=True)
            r = subprocess.run(cmd, cwd=tempfile.gettempdir(), stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                               timeout=150)
            files = glob.glob(os.path.join(out_dir, "**", name + "_counter_collection.csv"), recursive=True)
            if r.returncode != 0 and not files:
                return None, "rocprofv3 %s exited with %d" % (counter, r.returncode)
            per = {}
            for row in csv.DictReader(open(files[0])):
                per.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
            copy = sum(sum(v) for k, v in per.items() if "copyBuffer" in k)
            if copy <= 0:
                return None, "no copy dispatch in the %s run" % counter
            factors[name.split("_")[0]] = (COPY_MIB << 20) / (copy * 1024.0)
            for key, kname in KERNEL_NAMES.items():
                v = [x for k, xs in per.items() if re.search(kname, k) for x in xs]
                if v:
                    vals.setdefault(key, {})[name] = sum(v) / CHILD_FRAMES * 1024.0
    except subprocess.TimeoutExpired:
        return None, "rocprofv3 timed out"
    finally:
        shutil.rmtree(out_dir, ignore_errors=True)
    traffic = {k: v.get("fetch_size", 0.0) * factors["fetch"] + v.get("write_size", 0.0) * factors["write"]
               for k, v in vals.items()}
    return traffic, factors


def pmc_child(args):
    """The profiled child: one copy of known size, then two frames."""
    pk = importlib.import_module("imagej-elphel_amd")
    copy_calibration(COPY_MIB)
    sc = pk.synth.make_scene(W, H, seed=args.seed, disparity=("field", 0.0, 20.0), red_row=RED[0], red_col=RED[1])
    tp = make_processor(pk, sc)
    for _ in range(CHILD_FRAMES):
        run_step(tp)
    tp.sync()
    tp.close()


def make_processor(pk, sc):
    tp = pk.GPUTileProcessor(W, H, C, K, red_row=RED[0], red_col=RED[1])
    tp.setConvolutionKernels(sc.kernels, offsets=sc.offsets, force=True)
    tp.setBayerImages(sc.images, force=True)
    tp.setTasks(sc.tasks)
    tp.setCorrSigma(0.8)
    tp.setTexParams(keep_weights=0)
    tp.setAsync(True)
    return tp


def run_step(tp):
    """One frame: execConverCorrectTiles(0) and execCorrelate2D() as one call (one kernel per task
    block, the CLT handed from the conversion to the correlation in LDS and written once, for the
    textures), then textures and imclt_rbg in one pass over the converted tiles (the inverse CLT of
    each tile once for both outputs).  The per-tile LMA chain that correlate2D starts on its own
    stream overlaps the one-pass kernel and the next frame's conversion."""
    tp.execConvertCorrelate2D(0)
    tp.execTexturesImclt(True)


def batch_bench(pk, sc, frames=32, seed=7, repeats=3):
    """Batched form: `frames` quad frames, each with a random 5..100 % of the tiles as tasks (the
    frames share the synthetic images), one graph launch for convert+correlate2D+textures+imclt of
    all frames, against the same frames as sequential plain calls (setTasks + the four exec calls,
    each returning after the GPU finished).  Both include the task upload."""
    rng = np.random.default_rng(seed)
    T = len(sc.tasks)
    subsets = [np.sort(rng.choice(T, int(round(T * rng.uniform(0.05, 1.0))), replace=False)) for _ in range(frames)]
    total = sum(len(x) for x in subsets)
    # the frames' task lists are the caller's data, built before the timed regions of both paths
    frame_tasks = [np.ascontiguousarray(sc.tasks[sub]) for sub in subsets]
    # plain per-frame path
    tp = make_processor(pk, sc)
    tp.setAsync(False)
    best_plain = 1e30
    for _ in range(repeats):
        t0 = time.perf_counter()
        for ft in frame_tasks:
            tp.setTasks(ft)
            run_step(tp)
        best_plain = min(best_plain, time.perf_counter() - t0)
    tp.setTiming(True)
    tp.resetTiming()
    for ft in frame_tasks:
        tp.setTasks(ft)
        run_step(tp)
    plain_gpu = sum(tp.getTiming(k)[0] for k in ("convert_corr2d", "corr_lma", "tex_imclt"))
    tp.close()
    # batched
    bt = pk.GPUTileBatch(frames, W, H, C, K, red_row=RED[0], red_col=RED[1])
    bt.setConvolutionKernels(sc.kernels, offsets=sc.offsets, force=True)
    bt.setCorrSigma(0.8)
    bt.setTexParams(keep_weights=0)
    for f in range(frames):
        bt.setBayerImagesFrame(f, sc.images)
        bt.setTasksFrame(f, frame_tasks[f])
    tune = bt.autotune()
    bt.execBatch()   # graph capture
    bt.sync()
    best_batch = 1e30
    for _ in range(repeats):
        t0 = time.perf_counter()
        for f in range(frames):
            bt.setTasksFrame(f, frame_tasks[f])
        bt.execBatch()
        bt.sync()
        best_batch = min(best_batch, time.perf_counter() - t0)
    bt.setTiming(True)
    bt.resetTiming()
    bt.execBatch()
    batch_gpu = bt.getTiming("batch")[0]
    by_chunks = {}   # graph time of each pipeline depth the autotuner chooses from
    for chunks in (1, 2, 4, 8):
        bt.setBatchChunks(chunks)
        bt.execBatch()   # capture
        bt.resetTiming()
        bt.execBatch()
        by_chunks[chunks] = round(bt.getTiming("batch")[0], 3)
    bt.close()
    return {"frames": frames, "tasks": total, "coverage": "uniform 5..100 % per frame",
            "batched_tiles_per_s": round(total / best_batch, 1), "plain_tiles_per_s": round(total / best_plain, 1),
            "batched_ms": round(1e3 * best_batch, 3), "plain_ms": round(1e3 * best_plain, 3),
            "speedup": round(best_plain / best_batch, 3), "batched_gpu_ms": round(batch_gpu, 3),
            "plain_gpu_ms": round(plain_gpu, 3), "batched_gpu_ms_by_chunks": by_chunks, "tune": tune}


def configs_bench(pk, steps=50):
    """The other single-frame configurations, timed outside `value`: c0, the quad 512x512 Bayer
    RGGB frame (dense tasks, disparity 3 px: convert + correlate2D with the LMA chain + the one-pass
    textures/imclt_rbg), and c2, 16 monochrome 640x512 cameras on a circle (30 % of the tiles,
    all 120 pairs: convert + correlate2D + imclt_rbg); each beside the FP64 oracle of the same step
    on the host cores (one warm-up, then the median of 3 passes)."""
    import oracle as orc
    cores = host_cores()
    out = {}
    cases = (("c0", "quad 512x512 Bayer RGGB, dense, disparity 3 px",
              dict(width=512, height=512, seed=7, disparity=3.0, red_row=0, red_col=0), None),
             ("c2", "16 mono 640x512, 30 % tasks, 120 pairs",
              dict(width=640, height=512, num_cams=16, num_colors=1, seed=20261206, disparity=("field", 0.0, 6.0),
                   task_fraction=0.3), [(i, j, -1, 1) for i in range(16) for j in range(i + 1, 16)]))
    for key, name, kw, pairs in cases:
        sc = pk.synth.make_scene(**kw)
        bayer = sc.num_colors == 3
        tp = pk.GPUTileProcessor(sc.width, sc.height, sc.num_cams, sc.num_colors, red_row=sc.red_row,
                                 red_col=sc.red_col)
        tp.setConvolutionKernels(sc.kernels, offsets=sc.offsets, force=True)
        tp.setBayerImages(sc.images, force=True)
        tp.setTasks(sc.tasks)
        tp.setCorrSigma(0.8)
        if pairs:
            tp.setCorrPairs(pairs)
        if bayer:
            tp.setTexParams(keep_weights=0)
        tp.setAsync(True)

        def step():
            tp.execConverCorrectTiles(0)
            tp.execCorrelate2D()
            if bayer:
                tp.execTexturesImclt(True)
            else:
                tp.execImcltRbg(True)

        for _ in range(5):
            step()
        tp.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        tp.sync()
        ms = (time.perf_counter() - t0) / steps * 1e3
        tp.close()
        T = len(sc.tasks)
        g = orc.make_geom(sc.width, sc.height, sc.num_cams, sc.num_colors, sc.red_row, sc.red_col)
        valid = np.zeros((sc.height // 8) * (sc.width // 8), np.uint8)
        valid[((sc.tasks[:, 1] >> 16) * (sc.width // 8) + (sc.tasks[:, 1] & 0xFFFF))] = 1

        def cpu_step():
            clt = orc.convert_frame(g, sc.images, sc.tasks, kernels=sc.kernels, offsets=sc.offsets, nthreads=cores)
            orc.corr_frame(g, clt, sc.tasks, pairs=pairs, nthreads=cores)
            if bayer:
                orc.textures_frame(g, clt, sc.tasks, tex_params=orc.default_tex_params(keep_weights=0),
                                   nthreads=cores)
            orc.imclt_frame(g, clt, valid, lpf64=orc.lpf(0.8), nthreads=cores)

        cpu_step()
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            cpu_step()
            times.append(time.perf_counter() - t0)
        cpu_s = float(np.median(times))
        out[key] = {"workload": name, "tasks": T, "ms_per_step": round(ms, 4), "tiles_per_s": round(T / (ms * 1e-3), 1),
                    "cpu_port_fp64_tiles_per_s": round(T / cpu_s, 1), "cpu_cores": cores}
    return out


def adjoint_bench(pk, sc, repeats=5, seed=11):
    """Differentiable form: forward (convert + correlate2D) and the adjoint (port centres from the
    geometry, the fused convert/loss adjoint, the geometry adjoint and the LMA system reduction)
    per pass on the bench frame; then the extrinsic/disparity solve on a 2592x1936 frame rendered
    through the port geometry with extrinsics drawn N(0, 1e-3): 10 steps on the pairs'
    correlation-peak offsets (the residual form), and 5 such steps followed by 5 loss-adjoint LMA
    iterations.  Each reports the RMS error of the pairs' relative port offsets (what the
    correlations see) against the true geometry, from the library's own port geometry, in all and
    from the extrinsics alone (every tile's disparity re-fitted to the truth); the extrinsics
    themselves are determined only up to the quad's gauge (azimuths along rXY_x and tilts along
    rXY_y act as a common disparity)."""
    T = len(sc.tasks)
    tx, ty = sc.tasks[:, 1] & 0xFFFF, (sc.tasks[:, 1] >> 16) & 0xFFFF
    d = sc.disparity[ty, tx].astype(np.float32)
    words = np.full(T, 0xFF, np.int32)
    tp = pk.GPUTileProcessor(W, H, C, K, red_row=RED[0], red_col=RED[1])
    tp.setConvolutionKernels(sc.kernels, offsets=sc.offsets, force=True)
    tp.setBayerImages(sc.images, force=True)
    tp.setCorrSigma(0.8)
    tp.setPortGeometry(pk.GPUTileProcessor.defaultPortGeometry(W, H))
    tp.getTileSubcamOffsets(words, tx, ty, d)
    tp.execConverCorrectTiles(0)
    tp.execCorrelate2D()
    tp.getGeometryGradient()
    tp.setTiming(True)
    tp.resetTiming()
    for _ in range(repeats):
        tp.execConverCorrectTiles(0)
        tp.execCorrelate2D()
        tp.getGeometryGradient()
    fwd = (tp.getTiming("convert")[0] + tp.getTiming("corr2d")[0]) / repeats
    adj = tp.getTiming("geometry_adjoint")[0] / repeats
    tp.setTiming(False)
    # One gradient (13 extrinsics + every tile's disparity) by the adjoint, wall clock per call,
    # against the same gradient through the plain per-frame path: central differences of the loss,
    # one +/- pair of passes per extrinsic and one for all the disparities at once (each tile's loss
    # depends on its own disparity only), each pass the device port centres, convert and
    # correlate2D without the per-tile LMA.  (Reading the 28 losses back, a sum over the pairs'
    # correlation centres, is left out of the plain path's time.)
    tp.setAsync(True)
    tp.sync()
    t0 = time.perf_counter()
    for _ in range(repeats):
        tp.getGeometryGradient()
    grad_adj = 1e3 * (time.perf_counter() - t0) / repeats
    tp.setCorrLmaParams(enable=0)
    pg0 = list(pk.GPUTileProcessor.defaultPortGeometry(W, H).extrinsic)
    h_e, h_d = 1e-4, 1e-2

    def plain_pass(ext, disp):
        tp.setPortGeometry(pk.GPUTileProcessor.defaultPortGeometry(W, H, extrinsic=ext))
        tp.getTileSubcamOffsets(words, tx, ty, disp)
        tp.execConverCorrectTiles(0)
        tp.execCorrelate2D()

    plain_pass(pg0, d)
    tp.sync()
    t0 = time.perf_counter()
    for _ in range(repeats):
        for k in range(13):
            for sgn in (1.0, -1.0):
                e = list(pg0)
                e[k] += sgn * h_e
                plain_pass(e, d)
        for sgn in (1.0, -1.0):
            plain_pass(pg0, (d + sgn * h_d).astype(np.float32))
        tp.sync()
    grad_fd = 1e3 * (time.perf_counter() - t0) / repeats
    tp.close()
    # the solve on a frame rendered through the port geometry itself (camera images resampled at
    # the port positions of extrinsics drawn N(0, 1e-3) and a smooth disparity field), started
    # from zero extrinsics and the true disparities plus N(0, 0.05 px)
    rng = np.random.default_rng(seed)
    e_true = rng.normal(0, 1e-3, 13)
    print("bench: rendering the 2592x1936 geometry frame ...", file=sys.stderr, flush=True)
    gs = pk.synth.make_geometry_scene(W, H, e_true, seed=seed)
    inner = (gs.tx > 1) & (gs.tx < W // 8 - 2) & (gs.ty > 1) & (gs.ty < H // 8 - 2)
    tiles = np.nonzero(inner)[0]
    d_true = gs.disparity_tiles.ravel()[tiles].astype(np.float32)
    d0 = (d_true + rng.normal(0, 0.05, len(tiles))).astype(np.float32)
    gw, gx, gy = np.full(len(tiles), 0xFF, np.int32), gs.tx[tiles], gs.ty[tiles]
    gp = pk.GPUTileProcessor(W, H, 4, 3, kernels=False)
    gp.setBayerImages(gs.images, force=True)
    gp.setCorrSigma(0.8)
    pairs = [(0, 1), (2, 3), (0, 2), (1, 3), (0, 3), (2, 1)]

    def port_xy(extrinsic, disparity):   # the tasks' port centres from the library's port geometry
        gp.setPortGeometry(pk.GPUTileProcessor.defaultPortGeometry(W, H, extrinsic=list(extrinsic)))
        gp.getTileSubcamOffsets(gw, gx, gy, disparity)
        return gp.getTasks()[:, 2:].view(np.float32).reshape(len(tiles), 4, 2).astype(np.float64)

    true_xy = port_xy(e_true, d_true)

    def rel(xy):
        return np.stack([xy[:, i] - xy[:, j] for i, j in pairs], 1).reshape(len(xy), 12)

    true_rel = rel(true_xy)

    def errors(extrinsic, disparity):
        """RMS error of the pairs' relative port offsets against the true geometry, px: all of it,
        and what the extrinsics alone leave once each tile's disparity is re-fitted to the truth."""
        r = rel(port_xy(extrinsic, disparity)) - true_rel
        dd = np.asarray(disparity, np.float64).copy()
        for _ in range(2):
            r0 = rel(port_xy(extrinsic, dd)) - true_rel
            jac = (rel(port_xy(extrinsic, dd + 1e-2)) - rel(port_xy(extrinsic, dd))) / 1e-2
            dd -= (jac * r0).sum(1) / (jac * jac).sum(1)
        rx = rel(port_xy(extrinsic, dd)) - true_rel
        return float(np.sqrt(np.mean(r ** 2))), float(np.sqrt(np.mean(rx ** 2)))

    runs = {}
    start = errors(np.zeros(13), d0)
    for name, coarse in (("peak_offsets_10", 10), ("peak_offsets_5_lma_5", 5)):
        port_xy(np.zeros(13), d0)
        t0 = time.perf_counter()
        ext, disp, hist = gp.lmaExtrinsics(10, coarse_iter=coarse)
        ms = 1e3 * (time.perf_counter() - t0)
        end = errors(ext, disp)
        runs[name] = {"coarse_steps": coarse, "iterations": len(hist) - 1, "ms": round(ms, 2),
                      "rel_port_offset_px": [round(start[0], 4), round(end[0], 4)],
                      "from_extrinsics_px": [round(start[1], 4), round(end[1], 4)]}
    gp.close()
    return {"tasks": T, "forward_ms": round(fwd, 4), "adjoint_ms": round(adj, 4),
            "forward_tiles_per_s": round(T / (fwd * 1e-3), 1), "adjoint_tiles_per_s": round(T / (adj * 1e-3), 1),
            "gradient": {"adjoint_ms": round(grad_adj, 3), "plain_central_differences_ms": round(grad_fd, 3),
                         "plain_passes": 28, "speedup": round(grad_fd / grad_adj, 2)},
            "solve_tasks": len(tiles), "solve": runs}


def kernel_prep_bench(pk, seed=13, cams=4):
    """calculateCLTKernel on the device for the 2592x1936 kernel grid (164x123 kernels x 3 colours
    per camera, a 991 MB raw stack of 64x64 double-resolution PSFs): wall time per camera with the
    host-to-device upload, device time of the preparation kernels, and the FP64 oracle's time on
    the same stack on one host core."""
    import oracle
    tp = pk.GPUTileProcessor(W, H, C, K)
    kx, ky = tp.KERNELS_HOR, tp.KERNELS_VERT
    rng = np.random.default_rng(seed)
    raw = pk.synth.raw_kernel_stack(K, ky, kx, rng)
    sxy = np.stack([pk.synth.smooth_field(ky, kx, rng, -1.0, 1.0) for _ in range(2)], axis=-1).astype(np.float32)
    tp.calculateCLTKernel(0, raw, 64, sensor_xy=sxy)   # allocations
    tp.setTiming(True)
    tp.resetTiming()
    t0 = time.perf_counter()
    for cam in range(cams):
        tp.calculateCLTKernel(cam, raw, 64, sensor_xy=sxy)
    wall = (time.perf_counter() - t0) / cams
    dev_ms, launches = tp.getTiming("kernel_prep")
    dev_ms /= cams
    tp.close()
    n = kx * ky * K
    alg = raw.nbytes + n * (256 + 8) * 4 + n * 16 * 10   # raw read, kernels + extras written, centres
    t0 = time.perf_counter()
    oracle.prepare_kernels(raw, kx, ky, sensor_xy=sxy)
    cpu = time.perf_counter() - t0
    gbs = alg / (dev_ms * 1e-3) / 1e9
    return {"kernels_per_camera": n, "raw_MB": round(raw.nbytes / 1e6, 1), "wall_ms_per_camera": round(1e3 * wall, 2),
            "device_ms_per_camera": round(dev_ms, 4), "launches_per_camera": launches // cams,
            "alg_MB": round(alg / 1e6, 1), "GB_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
            "cpu_port_ms_per_camera": round(1e3 * cpu, 1), "cpu_cores": 1}


def rig_bench(pk, seed=17, steps=5, cpu_size=(512, 384)):
    """Two-quad rig (the correlation part of clt_bi_quad) on an 8-camera 2592x1936 frame: main
    quad rXY as the quad, the auxiliary quad at 2 baselines to the right with 0.75 of the baseline;
    per step convert (8 cameras) and clt_exec_corr2d_rig (both quads' correlate2D and tileCorrs
    LMA, the inter-camera correlation, its LMA and 2-D polynomial).  The FP64 oracle's rig stage is
    timed on all host cores on a smaller frame of the same kind."""
    import oracle as orc
    q = pk.geometry.QUAD_RXY
    rxy = np.concatenate([q, np.array([2.0, 0.0]) + 0.75 * q])
    sc = pk.synth.make_scene(W, H, num_cams=8, rxy=rxy, seed=seed, disparity=("field", 0.0, 10.0))
    tx, ty = sc.tasks[:, 1] & 0xFFFF, sc.tasks[:, 1] >> 16
    target = sc.disparity[ty, tx].astype(np.float32)
    tp = pk.GPUTileProcessor(W, H, 8, K)
    tp.setConvolutionKernels(sc.kernels, offsets=sc.offsets, force=True)
    tp.setBayerImages(sc.images, force=True)
    tp.setTasks(sc.tasks)
    tp.setCorrSigma(0.8)
    tp.execConverCorrectTiles(0)
    tp.execCorrelate2DRig(target)
    tp.setTiming(True)
    tp.resetTiming()
    t0 = time.perf_counter()
    for _ in range(steps):
        tp.execConverCorrectTiles(0)
        tp.execCorrelate2DRig(target)
    wall = (time.perf_counter() - t0) / steps
    rig_ms = tp.getTiming("rig")[0] / steps
    conv_ms = tp.getTiming("convert")[0] / steps
    b = tp.getCorr2DRig()["bimap"]
    tp.close()
    T = len(sc.tasks)
    cw, ch = cpu_size
    cs = pk.synth.make_scene(cw, ch, num_cams=8, rxy=rxy, seed=seed, disparity=("field", 0.0, 10.0))
    g = orc.make_geom(cw, ch, 8, K)
    clt = orc.convert_frame(g, cs.images, cs.tasks, kernels=cs.kernels, offsets=cs.offsets)
    cores = os.cpu_count() or 1
    t0 = time.perf_counter()
    n = 0
    while True:
        orc.rig_frame(g, clt, cs.tasks, nthreads=cores)
        n += 1
        dt = time.perf_counter() - t0
        if dt >= 3.0 or n >= 20:
            break
    return {"tasks": T, "cameras": 8, "ms_per_step": round(1e3 * wall, 3), "rig_ms": round(rig_ms, 4),
            "convert_ms": round(conv_ms, 4), "rig_tiles_per_s": round(T / (rig_ms * 1e-3), 1),
            "cross_fitted": int(np.isfinite(b[:, 11]).sum()),
            "cpu_port_rig_tiles_per_s": round(n * len(cs.tasks) / dt, 1), "cpu_cores": cores,
            "cpu_sample": "%dx%d 8-camera frame, %d tiles, the rig stage only, %d pass(es)" % (cw, ch, len(cs.tasks), n)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=20261204)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="run one device-to-device copy of MB MiB first (HBM counter calibration)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM counter runs")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-frames comparison")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/LMA timing")
    ap.add_argument("--no-kprep", action="store_true", help="skip the kernel preparation timing")
    ap.add_argument("--no-rig", action="store_true", help="skip the two-quad rig timing")
    ap.add_argument("--no-configs", action="store_true", help="skip the c0 / c2 configuration timings")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the tile path has no multi-GPU form yet); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    # the counter runs are child processes started before this process touches the GPU
    traffic, traffic_info = None, "skipped (--no-traffic)"
    if not args.no_traffic:
        traffic, traffic_info = pmc_traffic(args.seed)
        if traffic is None:
            print("bench: no HBM counters: %s" % traffic_info, file=sys.stderr, flush=True)
    pk = importlib.import_module("imagej-elphel_amd")
    if pk.load().clt_device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
    print("bench: generating the 2592x1936 quad frame ...", file=sys.stderr, flush=True)
    t = time.perf_counter()
    sc = pk.synth.make_scene(W, H, seed=args.seed, disparity=("field", 0.0, 20.0), red_row=RED[0], red_col=RED[1])
    print("bench: scene ready in %.1f s, %d tiles" % (time.perf_counter() - t, len(sc.tasks)), file=sys.stderr, flush=True)

    tp = make_processor(pk, sc)

    def step():
        run_step(tp)

    for _ in range(args.warmup):
        step()
    tp.sync()
    tp.setTiming(True, reserve=args.steps)   # the events exist before the timed region
    tp.resetTiming()
    tp.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    tp.sync()
    dt = time.perf_counter() - t0
    ms_step = 1000 * dt / args.steps
    T = len(sc.tasks)
    value = T / (dt / args.steps)

    ab = algorithmic_bytes(sc, planes=4, n_pairs=6, strip_h=9)
    kernels = {}
    for name in STEP_KERNELS:
        ms, cnt = tp.getTiming(name)
        avg = ms / max(cnt, 1)
        gbs = ab[name] / (avg * 1e-3) / 1e9
        kernels[name] = {"avg_ms": round(avg, 4), "launches": cnt, "alg_MB": round(ab[name] / 1e6, 1),
                         "GB_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    total_alg = sum(ab[k] for k in STEP_KERNELS)
    # the per-tile LMA chain (FP64 fits, data-dependent iteration counts): its time, not a byte rate
    ms, cnt = tp.getTiming("corr_lma")
    lma_ms = ms / max(cnt, 1)
    kernels["corr_lma"] = {"avg_ms": round(lma_ms, 4), "launches": cnt, "ns_per_tile": round(lma_ms * 1e6 / T, 2),
                           "bound": "fp64 valu"}
    roof = {"bound": "hbm", "kernel": dom, "achieved": kernels[dom]["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": kernels[dom]["frac"], "traffic": None,
            "frame_alg_MB": round(total_alg / 1e6, 1),
            "frame_frac": round(total_alg / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    # The timed steps run the LMA concurrently with textures/imclt and convert, which stretches their
    # event-timed durations; a few steps more with every call waiting for its kernels give each
    # kernel's own duration (not part of `value`).
    tp.setAsync(False)
    tp.resetTiming()
    for _ in range(3):
        step()
    isolated = {}
    for name in STEP_KERNELS + ("corr_lma",):
        ms, cnt = tp.getTiming(name)
        avg = ms / max(cnt, 1)
        isolated[name] = {"avg_ms": round(avg, 4)}
        if name in ab:
            gbs = ab[name] / (avg * 1e-3) / 1e9
            isolated[name].update({"GB_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
    roof["isolated_frac"] = isolated[dom].get("frac")
    gpu_out = gpu_outputs(tp, sc)
    tp.close()
    batch = None
    if not args.no_batch:
        print("bench: batched frames ...", file=sys.stderr, flush=True)
        batch = batch_bench(pk, sc)
    adjoint = None
    if not args.no_adjoint:
        print("bench: adjoint and LMA ...", file=sys.stderr, flush=True)
        adjoint = adjoint_bench(pk, sc)
    kprep = None
    if not args.no_kprep:
        print("bench: kernel preparation ...", file=sys.stderr, flush=True)
        kprep = kernel_prep_bench(pk)
    configs = None
    if not args.no_configs:
        print("bench: c0 and c2 configurations ...", file=sys.stderr, flush=True)
        configs = configs_bench(pk)
    rig = None
    if not args.no_rig:
        print("bench: two-quad rig ...", file=sys.stderr, flush=True)
        rig = rig_bench(pk)
    if traffic is None:
        roof["traffic_note"] = traffic_info
    else:
        for k, v in traffic.items():
            kernels[k]["traffic_MB"] = round(v / 1e6, 1)
        if dom in traffic:
            roof["traffic"] = round(traffic[dom])
        roof["counter_scale"] = {k: round(v, 3) for k, v in traffic_info.items()}
    cpu, errors = None, None
    if not args.no_cpu:
        print("bench: timing the CPU baseline ...", file=sys.stderr, flush=True)
        cpu, ref = cpu_baseline(sc)
        errors = {"vs": "the FP64 oracle on the same frame", "gpu": output_errors(gpu_out, ref[64]),
                  "oracle_fp32": output_errors(ref[32], ref[64])}
        del ref
    del gpu_out
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "tiles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "quad 2592x1936 Bayer RGGB frame, 78408 tiles, dense tasks, disparity 0-20 px; "
                               "convert+correlate2D(6 pairs, per-tile LMA chain)+textures(RGBA)+imclt_rbg "
                               "(one pass)",
                   "tiles": T, "cameras": C, "colors": K, "pairs": 6},
        "roofline": roof, "kernels": kernels, "kernels_isolated": isolated, "cpu_baseline": cpu, "errors": errors,
        "batch": batch, "adjoint": adjoint,
        "kernel_prep": kprep, "rig": rig, "configs": configs,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
