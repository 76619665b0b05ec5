"""GPUTileProcessor: Python mirror of the reference's JCuda wrapper over the C ABI.

Same method names, argument meaning and call order as
src/main/java/com/elphel/imagej/gpu/GPUTileProcessor.java (GTP), backed by libclt_hip.so.
Methods without a reference counterpart (correlation, textures, timing) are named in the same
style.  Errors of the C ABI raise ``CltError`` (the reference raises ``CudaException``).
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import CltError


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class TpTask:
    """GTP:132-166 TpTask: task word, tile (tx, ty), target disparity, port centres xy[C][2]."""

    def __init__(self, tx, ty, target_disparity=0.0, task=1, xy=None):
        self.tx, self.ty = tx, ty
        self.target_disparity = target_disparity
        self.task = task
        self.xy = xy

    def as_record(self):
        xy = np.asarray(self.xy, np.float32)
        rec = np.zeros(2 + xy.size, np.int32)
        rec[0] = self.task
        rec[1] = self.tx | (self.ty << 16)
        rec[2:] = xy.reshape(-1).view(np.int32)
        return rec


class GPUTileProcessor:
    """GPUTileProcessor() (GTP:243-360) for one frame geometry.

    Defaults are the reference statics (GTP:80-97): 2592x1936, 4 cameras, 3 colours,
    164x123 kernels, Bayer phase GR/BG (red at odd columns of even rows, TPC:113-115)."""

    DTT_SIZE = 8

    def __init__(self, width=2592, height=1936, num_cams=4, num_colors=3, red_row=0, red_col=1,
                 kernels=True, kernel_step=16, max_tasks=0, _frames=0):
        self.lib = _abi.load()
        g = _abi.clt_geom(width=width, height=height, num_cams=num_cams, num_colors=num_colors,
                          red_row=red_row, red_col=red_col,
                          kernels_hor=(width // kernel_step + 2) if kernels else 0,
                          kernels_vert=(height // kernel_step + 2) if kernels else 0,
                          kernel_step=kernel_step, max_tasks=max_tasks)
        self.geom = g
        self.IMG_WIDTH, self.IMG_HEIGHT, self.NUM_CAMS, self.NUM_COLORS = width, height, num_cams, num_colors
        self.KERNELS_HOR, self.KERNELS_VERT = g.kernels_hor, g.kernels_vert
        self.tiles_x, self.tiles_y = width // 8, height // 8
        h = ctypes.c_void_p()
        if _frames:
            rc = self.lib.clt_create_batch(ctypes.byref(g), _frames, ctypes.byref(h))
        else:
            rc = self.lib.clt_create(ctypes.byref(g), ctypes.byref(h))
        if rc != 0:
            raise CltError(rc, "clt_create failed (geometry %dx%d, %d cameras)" % (width, height, num_cams))
        self.h = h
        self.kernels_set = False
        self.bayer_set = False
        self.num_task_tiles = 0
        self.num_pairs = 6 if num_cams == 4 else 0
        self.strip_height = 9
        self.keep_weights = True

    # ------------------------------------------------------------------ plumbing
    def _chk(self, rc):
        if rc != 0:
            raise CltError(rc, self.lib.clt_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.clt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------------ inputs (GTP:363-527)
    def setTasks(self, tile_tasks, use_aux=False):
        """GTP:363-371.  tile_tasks: int32 records [n][2+2C] or a list of TpTask."""
        if isinstance(tile_tasks, (list, tuple)):
            recs = np.stack([t.as_record() for t in tile_tasks]) if tile_tasks else np.zeros((0, 2 + 2 * self.NUM_CAMS), np.int32)
        else:
            recs = np.ascontiguousarray(tile_tasks, np.int32)
        if recs.ndim != 2 or recs.shape[1] != 2 + 2 * self.NUM_CAMS:
            raise ValueError("task records must be [n][%d]" % (2 + 2 * self.NUM_CAMS))
        self._chk(self.lib.clt_set_tasks(self.h, _ptr(recs), recs.shape[0]))
        self.num_task_tiles = recs.shape[0]
        self._tasks = recs

    # ------------------------------------------------------------ port geometry (GTP:686-778)
    @staticmethod
    def defaultPortGeometry(width=2592, height=1936, **kw):
        """clt_port_geometry with the reference defaults (zero distortion/extrinsics, GC:37-78);
        keyword arguments set fields (distortion, extrinsic, roll_deg, rxy, pxy0 as sequences)."""
        p = _abi.clt_port_geometry()
        _abi.load().clt_default_port_geometry(ctypes.byref(p), width, height)
        for k, v in kw.items():
            if k in ("distortion", "extrinsic", "roll_deg"):
                arr = getattr(p, k)
                for i, x in enumerate(v):
                    arr[i] = x
            elif k in ("rxy", "pxy0"):
                arr = getattr(p, k)
                for i, (x, y) in enumerate(v):
                    arr[i][0], arr[i][1] = x, y
            else:
                setattr(p, k, v)
        return p

    def setPortGeometry(self, geometry):
        """The GeometryCorrection of the camera (rotation matrices, reverse distortion table)."""
        self._chk(self.lib.clt_set_port_geometry(self.h, ctypes.byref(geometry)))

    @staticmethod
    def rigOffset(baseline=1256.0, main_disparity_radius=None, aux_disparity_radius=None, **kw):
        """clt_rig_offset (GeometryCorrection.RigOffset GC:256-290); the disparity radii default to
        sqrt(2) * 105 mm (GC:81, GC:1757)."""
        r = _abi.clt_rig_offset(baseline=baseline)
        dflt = 2 ** 0.5 * 105.0
        r.main_disparity_radius = dflt if main_disparity_radius is None else main_disparity_radius
        r.aux_disparity_radius = dflt if aux_disparity_radius is None else aux_disparity_radius
        for k, v in kw.items():
            setattr(r, k, v)
        return r

    def setPortGeometryRig(self, main, aux, rig_offset):
        """The two quads of a rig (an 8-camera processor): main ports 0..3, auxiliary 4..7 with the
        rig offsets, as clt_bi_quad computes them (ID:8970-8993)."""
        self._chk(self.lib.clt_set_port_geometry_rig(self.h, ctypes.byref(main), ctypes.byref(aux),
                                                     ctypes.byref(rig_offset)))

    def getTileSubcamOffsets(self, task, tx, ty, target_disparity):
        """GTP:686-778 followed by setTasks: the port centres of every task are computed on the
        GPU from the port geometry.  task, tx, ty, target_disparity: per-task arrays."""
        task = np.ascontiguousarray(task, np.int32)
        txy = np.ascontiguousarray(np.asarray(tx, np.int32) | (np.asarray(ty, np.int32) << 16), np.int32)
        disp = np.ascontiguousarray(target_disparity, np.float32)
        n = len(task)
        if txy.shape != (n,) or disp.shape != (n,):
            raise ValueError("task, tx, ty and target_disparity must have one entry per task")
        self._chk(self.lib.clt_set_tasks_geometry(self.h, _ptr(task), _ptr(txy), _ptr(disp), n))
        self.num_task_tiles = n
        self._tasks = None

    def setFullFrameImages(self, target_disparity, task=0xFF):
        """GTP:472-527: one task per tile of the frame at the given disparity (a scalar or
        [tilesY * tilesX] values), port centres from the port geometry (setPortGeometry or
        setPortGeometryRig) on the device, then set as the task list.  The reference writes the task
        word 1 (any non-zero word converts a tile); here bits 0..7 also select the ports the textures
        combine, so the default uses all of them.  Returns the task records."""
        tiles_x, tiles_y = self.IMG_WIDTH // self.DTT_SIZE, self.IMG_HEIGHT // self.DTT_SIZE
        ty, tx = np.mgrid[0:tiles_y, 0:tiles_x]
        d = np.broadcast_to(np.asarray(target_disparity, np.float32), (tiles_y * tiles_x,))
        words = np.full(tiles_y * tiles_x, task, np.int32)
        self.getTileSubcamOffsets(words, tx.ravel(), ty.ravel(), d)
        return self.getTasks()

    def getTasks(self):
        """The current task records [n][2+2C] (int32; port centres as float bits)."""
        out = np.zeros((self.num_task_tiles, 2 + 2 * self.NUM_CAMS), np.int32)
        self._chk(self.lib.clt_get_tasks(self.h, _ptr(out)))
        return out

    def getPortDerivatives(self):
        """[n][8][14] derivatives of the port centres (x0, y0, ..., x3, y3) by the 13 extrinsic
        parameters (GC:2245-2306) and by the disparity (last column)."""
        out = np.zeros((self.num_task_tiles, 8, 14), np.float64)
        self._chk(self.lib.clt_get_port_derivatives(self.h, _ptr(out)))
        return out

    def setConvolutionKernel(self, kernel, kernel_offsets, ncam):
        """GTP:373-379: float kernels [KY][KX][K][4][8][8] ([v][h]) and offsets [KY][KX][K][8]."""
        k = np.ascontiguousarray(kernel, np.float32)
        o = np.ascontiguousarray(kernel_offsets, np.float32)
        nk = self.KERNELS_HOR * self.KERNELS_VERT * self.NUM_COLORS
        if k.size != nk * 256 or o.size != nk * 8:
            raise ValueError("kernel arrays do not match the %dx%d kernel grid" % (self.KERNELS_HOR, self.KERNELS_VERT))
        self._chk(self.lib.clt_set_kernels(self.h, ncam, _ptr(k), _ptr(o)))

    def setConvolutionKernels(self, clt_kernels, force=False, offsets=None):
        """GTP:381-435.  clt_kernels: per camera arrays in upload layout (as produced by
        synth.clt_kernels) with ``offsets`` given separately."""
        if self.kernels_set and not force:
            return
        for ncam in range(self.NUM_CAMS):
            self.setConvolutionKernel(clt_kernels[ncam], offsets[ncam], ncam)
        self.kernels_set = True

    def calculateCLTKernel(self, ncam, kernel_stack, kernel_size=64, sensor_xy=None, norm_kern=True):
        """QC:433-668 for one camera, on the device, installing the result as setConvolutionKernel
        does.  kernel_stack: [NUM_COLORS][KERNELS_VERT*kernel_size][KERNELS_HOR*kernel_size] raw
        double-resolution kernels; sensor_xy: [KERNELS_VERT][KERNELS_HOR][2] sensor correction at
        the kernel centres (QC:603-610) or None."""
        raw = np.ascontiguousarray(kernel_stack, np.float32)
        shape = (self.NUM_COLORS, self.KERNELS_VERT * kernel_size, self.KERNELS_HOR * kernel_size)
        if raw.shape != shape:
            raise ValueError("kernel stack %s does not match %s" % (raw.shape, shape))
        sxy = None
        if sensor_xy is not None:
            sxy = np.ascontiguousarray(sensor_xy, np.float32)
            if sxy.size != self.KERNELS_VERT * self.KERNELS_HOR * 2:
                raise ValueError("sensor_xy does not match the %dx%d kernel grid" % (self.KERNELS_HOR, self.KERNELS_VERT))
        self._chk(self.lib.clt_prepare_kernels(self.h, ncam, _ptr(raw), kernel_size,
                                               None if sxy is None else _ptr(sxy), 1 if norm_kern else 0))

    def getConvolutionKernel(self, ncam):
        """The installed kernels of one camera: (kernels [KY][KX][K][4][8][8] ([v][h]),
        offsets [KY][KX][K][8]), the layout setConvolutionKernel takes."""
        shp = (self.KERNELS_VERT, self.KERNELS_HOR, self.NUM_COLORS)
        k = np.empty(shp + (4, 8, 8), np.float32)
        o = np.empty(shp + (8,), np.float32)
        self._chk(self.lib.clt_get_kernels(self.h, ncam, _ptr(k), _ptr(o)))
        return k, o

    def clearConvolutionKernels(self):
        self._chk(self.lib.clt_clear_kernels(self.h))
        self.kernels_set = False

    def setBayerImage(self, bayer_image, ncam):
        """GTP:437-450: one mosaic [H][W] (or flat) float image."""
        img = np.ascontiguousarray(bayer_image, np.float32).reshape(self.IMG_HEIGHT, -1)
        self._chk(self.lib.clt_set_image(self.h, ncam, _ptr(img), img.shape[1]))

    def setBayerImages(self, bayer_data, force=False):
        """GTP:452-468: bayer_data [C][3][H*W] colour planes (summed into mosaics) or [C][H][W]."""
        if self.bayer_set and not force:
            return
        for ncam in range(self.NUM_CAMS):
            d = np.asarray(bayer_data[ncam])
            if d.ndim == 2 and d.shape[0] == self.NUM_COLORS and self.NUM_COLORS == 3:
                d = d.sum(axis=0)
            self.setBayerImage(d, ncam)
        self.bayer_set = True

    def setLpfRbg(self, sigma_r, sigma_b, sigma_g):
        """GTP:780-812: convert-stage LPF per colour."""
        self._chk(self.lib.clt_set_lpf_rbg(self.h, sigma_r, sigma_b, sigma_g))

    @staticmethod
    def setCltLpf(sigma):
        """GTP:814-848 setCltLpf followed by dttt_iiie * 2n: 64 multipliers ([h][v])."""
        out = np.zeros(64, np.float32)
        rc = _abi.load().clt_lpf_coeffs(sigma, _ptr(out))
        if rc != 0:
            raise CltError(rc, "clt_lpf_coeffs")
        return out

    def setCorrSigma(self, sigma):
        """corr_sigma of ImageDtt (ID:1662-1690): LPF of correlation, textures and imclt."""
        self._chk(self.lib.clt_set_corr_sigma(self.h, sigma))

    def setCorrPairs(self, pairs):
        """pairs: [(cam1, cam2, dir, scale)], dir -1 for pairs off the hor/vert/diagonal axes."""
        arr = (_abi.clt_pair * len(pairs))(*[_abi.clt_pair(*p) for p in pairs])
        self._chk(self.lib.clt_set_pairs(self.h, arr, len(pairs)))
        self.num_pairs = len(pairs)

    def setCorrParams(self, **kw):
        p = _abi.clt_corr_params()
        self.lib.clt_default_corr_params(ctypes.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        self._chk(self.lib.clt_set_corr_params(self.h, ctypes.byref(p)))
        self.strip_height = p.strip_height

    def setCorrLmaParams(self, **kw):
        """Per-tile correlation LMA and foreground correction (ImageDttParameters IDP:32-110);
        reference defaults (enabled) updated by the keywords, e.g. enable=0."""
        p = _abi.clt_corr_lma_params()
        self.lib.clt_default_corr_lma_params(ctypes.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        self._chk(self.lib.clt_set_corr_lma_params(self.h, ctypes.byref(p)))

    def setRigParams(self, **kw):
        """Two-quad rig (BiQuadParameters BQP:32-83, IDP:81-112); reference defaults updated by
        the keywords, e.g. lt_rad=1, notch_mode=1, use_xy_poly=0."""
        p = _abi.clt_rig_params()
        self.lib.clt_default_rig_params(ctypes.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        self._chk(self.lib.clt_set_rig_params(self.h, ctypes.byref(p)))

    def setTexParams(self, port_offsets=None, **kw):
        p = _abi.clt_tex_params()
        self.lib.clt_default_tex_params(ctypes.byref(p), self.NUM_CAMS)
        if port_offsets is not None:
            for i, (x, y) in enumerate(port_offsets):
                p.port_offsets[i][0], p.port_offsets[i][1] = x, y
        for k, v in kw.items():
            setattr(p, k, v)
        self._chk(self.lib.clt_set_tex_params(self.h, ctypes.byref(p)))
        self.keep_weights = bool(p.keep_weights)

    # ------------------------------------------------------------------ kernels
    def execConverCorrectTiles(self, lpf_mask=0):
        """GTP:530-558 (the reference passes lpf_mask 7; the CPU path applies no LPF here)."""
        self._chk(self.lib.clt_exec_convert(self.h, lpf_mask))

    def execImcltRbg(self, apply_lpf=False):
        """GTP:560-597."""
        self._chk(self.lib.clt_exec_imclt(self.h, int(apply_lpf)))

    def execCorrelate2D(self):
        self._chk(self.lib.clt_exec_corr2d(self.h))

    def execConvertCorrelate2D(self, lpf_mask=0):
        """execConverCorrectTiles(lpf_mask) then execCorrelate2D() with the same outputs; in the quad
        layout one kernel that hands each task's CLT from the conversion to the correlation on chip."""
        self._chk(self.lib.clt_exec_convert_corr2d(self.h, lpf_mask))

    def execTextures(self, extra_disparity=None):
        ed = None if extra_disparity is None else np.ascontiguousarray(extra_disparity, np.float32)
        self._chk(self.lib.clt_exec_textures(self.h, _ptr(ed)))

    def execTexturesImclt(self, apply_lpf=True, extra_disparity=None):
        """execTextures then execImcltRbg (ID:2662-2812, GTP:560-597) with the same outputs; one pass
        over the converted tiles when there is no extra disparity, apply_lpf is set and every task
        has its own tile and non-zero flags."""
        ed = None if extra_disparity is None else np.ascontiguousarray(extra_disparity, np.float32)
        self._chk(self.lib.clt_exec_textures_imclt(self.h, _ptr(ed), int(apply_lpf)))

    def execCombineTextures(self, overlap=True, sharp_alpha=False):
        """combineRGBATiles (ID:3855-3991) of the last textures: overlap-added planes, or with
        overlap=False the (2H x 2W) mosaic of whole tiles; sharp_alpha: alpha from tile centres."""
        self._chk(self.lib.clt_exec_combine_rgba_tiles(self.h, int(bool(overlap)), int(bool(sharp_alpha))))
        self._combine_scale = 1 if overlap else 2

    # ------------------------------------------------------------------ outputs
    def getRBG(self, ncam):
        """GTP:599-626: [colors][(H+8)*(W+8)] float32."""
        out = np.empty((self.NUM_COLORS, (self.IMG_HEIGHT + 8) * (self.IMG_WIDTH + 8)), np.float32)
        self._chk(self.lib.clt_get_rbg(self.h, ncam, _ptr(out)))
        return out

    def getCLT(self, ncam):
        """[tilesY*tilesX][colors][4][8][8] float32 ([v][h] blocks)."""
        out = np.empty((self.tiles_x * self.tiles_y, self.NUM_COLORS, 4, 8, 8), np.float32)
        self._chk(self.lib.clt_get_clt(self.h, ncam, _ptr(out)))
        return out

    def getCorr2D(self):
        """(corr [n][P][15][15], combined strip [n][h][15], disparity [n][DISP_LAYERS])."""
        n = self.num_task_tiles
        corr = np.empty((n, self.num_pairs, 15, 15), np.float32)
        combo = np.empty((n, self.strip_height, 15), np.float32)
        disp = np.empty((n, _abi.DISP_LAYERS), np.float32)
        self._chk(self.lib.clt_get_corr2d(self.h, _ptr(corr), _ptr(combo), _ptr(disp)))
        return corr, combo, disp

    def getCorrPeakOffsets(self, radius=3.9):
        """Per task and quad pair (dx, dy, correlation at the maximum): the centre of mass of the
        positive correlation within radius of the pair's integer maximum, pixels from the centre
        (the clt_mismatch measurement, ID:2342-2420, as the coarse extrinsic stage uses it)."""
        out = np.empty((self.num_task_tiles, 6, 3), np.float32)
        self._chk(self.lib.clt_get_corr_offsets(self.h, float(radius), _ptr(out)))
        return out

    def execCorrelate2DRig(self, target_disparity=None):
        """clt_bi_quad (ID:8731-9504) correlation part on an 8-camera handle (main quad 0..3,
        auxiliary 4..7); target_disparity: per task (the BI_TARGET layer) or None."""
        tg = None
        if target_disparity is not None:
            tg = np.ascontiguousarray(np.broadcast_to(np.asarray(target_disparity, np.float32), (self.num_task_tiles,)))
        self._chk(self.lib.clt_exec_corr2d_rig(self.h, None if tg is None else _ptr(tg)))

    def getCorr2DRig(self):
        """dict: bimap [n][BI_LAYERS] (layers named by _abi.BI_NAMES), inter [n][15][15],
        corr_main / corr_aux [n][6][15][15], disp_main / disp_aux [n][DISP_LAYERS]."""
        n = self.num_task_tiles
        out = {"bimap": np.empty((n, len(_abi.BI_NAMES)), np.float32), "inter": np.empty((n, 15, 15), np.float32),
               "corr_main": np.empty((n, 6, 15, 15), np.float32), "corr_aux": np.empty((n, 6, 15, 15), np.float32),
               "disp_main": np.empty((n, _abi.DISP_LAYERS), np.float32),
               "disp_aux": np.empty((n, _abi.DISP_LAYERS), np.float32)}
        self._chk(self.lib.clt_get_corr2d_rig(self.h, *[_ptr(out[k]) for k in ("bimap", "inter", "corr_main", "corr_aux",
                                                                             "disp_main", "disp_aux")]))
        return out

    def execTexturesRig(self):
        """Both quads' textures after execCorrelate2DRig (ID:9360-9398)."""
        self._chk(self.lib.clt_exec_textures_rig(self.h))

    def getTexturesRig(self):
        """((tiles_main, tiles_aux) [n][planes][16][16], (stats_main, stats_aux) [n][16]); planes
        12 with keep_weights, else 4."""
        n = self.num_task_tiles
        planes = 12 if self.keep_weights else 4
        t = [np.empty((n, planes, 16, 16), np.float32) for _ in range(2)]
        st = [np.empty((n, 16), np.float32) for _ in range(2)]
        self._chk(self.lib.clt_get_textures_rig(self.h, _ptr(t[0]), _ptr(t[1]), _ptr(st[0]), _ptr(st[1])))
        return tuple(t), tuple(st)

    def getTextures(self):
        """(tiles [n][planes][16][16], stats [n][4C])."""
        n = self.num_task_tiles
        planes = (4 + self.NUM_CAMS + 4) if self.keep_weights else 4
        tiles = np.empty((n, planes, 16, 16), np.float32)
        stats = np.empty((n, 4 * self.NUM_CAMS), np.float32)
        self._chk(self.lib.clt_get_textures(self.h, _ptr(tiles), _ptr(stats)))
        return tiles, stats

    def getCombinedTextures(self):
        planes = (4 + self.NUM_CAMS + 4) if self.keep_weights else 4
        k = getattr(self, "_combine_scale", 1)
        out = np.empty((planes, k * self.IMG_HEIGHT, k * self.IMG_WIDTH), np.float32)
        self._chk(self.lib.clt_get_combined_textures(self.h, _ptr(out)))
        return out

    # ------------------------------------------------------------------ correlation-peak loss adjoint
    def execCorrAdjoint(self):
        """Loss -sum_p C_p(0) of every task and its gradient by the port centres (quad, fat zero 0)."""
        self._chk(self.lib.clt_exec_corr_adjoint(self.h))

    def getCorrAdjoint(self):
        """dict of arrays: loss [n], c [n][6], kx [n][6], ky [n][6], dxy [n][4][2]."""
        out = np.empty((self.num_task_tiles, _abi.CORR_ADJ), np.float32)
        self._chk(self.lib.clt_get_corr_adjoint(self.h, _ptr(out)))
        return {"loss": out[:, 0], "c": out[:, 1:7], "kx": out[:, 7:13], "ky": out[:, 13:19],
                "dxy": out[:, 19:27].reshape(-1, 4, 2)}

    def getGeometryGradient(self):
        """(dl/d extrinsics [13] float64, dl/d disparity [n] float32, loss) of the tasks set by
        getTileSubcamOffsets."""
        ge = np.zeros(13, np.float64)
        gd = np.zeros(self.num_task_tiles, np.float32)
        loss = ctypes.c_double()
        self._chk(self.lib.clt_exec_geometry_adjoint(self.h, _ptr(ge), _ptr(gd), ctypes.byref(loss)))
        return ge, gd, loss.value

    def getGeometryGradientNoAlloc(self):
        """getGeometryGradient through clt_exec_geometry_adjoint_noalloc: no device allocation (the
        buffers are the library's static scratch, at most MAX_TASKS_NOALLOC tasks)."""
        ge = np.zeros(13, np.float64)
        gd = np.zeros(self.num_task_tiles, np.float32)
        loss = ctypes.c_double()
        self._chk(self.lib.clt_exec_geometry_adjoint_noalloc(self.h, _ptr(ge), _ptr(gd), ctypes.byref(loss)))
        return ge, gd, loss.value

    MAX_TASKS_NOALLOC = 80000   # CLT_NOALLOC_MAX_TASKS

    def lmaExtrinsics(self, num_iter=10, adjust_disparity=True, **kw):
        """LMA of the extrinsics (and disparities) on the correlation-peak loss; returns
        (extrinsics [13], disparities [n], loss history)."""
        p = _abi.clt_lma_params()
        self.lib.clt_default_lma_params(ctypes.byref(p))
        p.num_iter = num_iter
        p.adjust_disparity = int(adjust_disparity)
        for k, v in kw.items():
            setattr(p, "lambda_" if k == "lambda" else k, v)
        ext = np.zeros(13, np.float64)
        disp = np.zeros(self.num_task_tiles, np.float32)
        hist = np.full(num_iter + 1, np.nan)
        its = ctypes.c_int32()
        self._chk(self.lib.clt_lma_extrinsics(self.h, ctypes.byref(p), _ptr(ext), _ptr(disp), _ptr(hist),
                                              ctypes.byref(its)))
        return ext, disp, hist[:its.value + 1]

    # ------------------------------------------------------------------ execution control
    TUNE_KNOBS = ("corr_tiles_per_block", "imclt_rows", "chunks", "tex_imclt_rows", "tex_lds_layout")

    def getTune(self):
        """Launch shapes in use (clt_get_tune_all): correlate2D tiles per workgroup, imclt rows per
        segment, batch frame chunks, one-pass textures/imclt rows per segment and LDS tile layout."""
        n = len(self.TUNE_KNOBS)
        v = (ctypes.c_int32 * n)()
        self._chk(self.lib.clt_get_tune_all(self.h, v, n))
        return dict(zip(self.TUNE_KNOBS, v))

    def setTune(self, **knobs):
        """Pins launch shapes (clt_set_tune_all), e.g. the result of GPUTileBatch.autotune()."""
        unknown = set(knobs) - set(self.TUNE_KNOBS)
        if unknown:
            raise ValueError("unknown tuning knobs: %s" % sorted(unknown))
        cur = self.getTune()
        cur.update(knobs)
        v = (ctypes.c_int32 * len(self.TUNE_KNOBS))(*[int(cur[k]) for k in self.TUNE_KNOBS])
        self._chk(self.lib.clt_set_tune_all(self.h, v, len(self.TUNE_KNOBS)))

    def setAsync(self, on):
        self._chk(self.lib.clt_set_async(self.h, int(on)))

    def sync(self):
        self._chk(self.lib.clt_sync(self.h))

    def setTiming(self, on, reserve=0):
        """HIP events around each kernel; reserve: create the events of that many launches of each
        kernel now rather than at first use."""
        self._chk(self.lib.clt_set_timing(self.h, max(int(bool(on)), int(reserve) if on else 0)))

    def resetTiming(self):
        self._chk(self.lib.clt_reset_timing(self.h))

    def getTiming(self, kernel):
        ms = ctypes.c_double()
        cnt = ctypes.c_int32()
        self._chk(self.lib.clt_get_timing(self.h, _abi.TIMING_KERNELS[kernel], ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value


class GPUTileBatch(GPUTileProcessor):
    """Many frames (camera sets of one geometry) with uneven task lists, processed by one
    hipGraph launch per call (clt_exec_batch); parameters and kernels are shared."""

    CONVERT, CORR2D, TEXTURES, IMCLT = 1, 2, 4, 8

    def __init__(self, num_frames, width=2592, height=1936, num_cams=4, num_colors=3, red_row=0, red_col=1,
                 kernels=True, kernel_step=16, max_tasks=0):
        super().__init__(width, height, num_cams, num_colors, red_row, red_col, kernels, kernel_step,
                         max_tasks, _frames=num_frames)
        self.num_frames = num_frames
        self.frame_tasks = [0] * num_frames

    def setBayerImagesFrame(self, frame, images):
        """images: [C][H][W] float mosaics of one frame."""
        for c in range(self.NUM_CAMS):
            img = np.ascontiguousarray(images[c], np.float32).reshape(self.IMG_HEIGHT, -1)
            self._chk(self.lib.clt_set_image_frame(self.h, frame, c, _ptr(img), img.shape[1]))

    def setTasksFrame(self, frame, recs):
        recs = np.ascontiguousarray(recs, np.int32)
        self._chk(self.lib.clt_set_tasks_frame(self.h, frame, _ptr(recs), recs.shape[0]))
        self.frame_tasks[frame] = recs.shape[0]

    def autotune(self):
        self._chk(self.lib.clt_autotune(self.h))
        return self.getTune()

    def setBatchChunks(self, chunks):
        """Frame chunks of the pipelined batch (clt_autotune picks one of 1, 2, 4, 8)."""
        self._chk(self.lib.clt_set_batch_chunks(self.h, int(chunks)))

    def execBatch(self, stages=15):
        self._chk(self.lib.clt_exec_batch(self.h, stages))

    def getCorr2DFrame(self, frame):
        n = self.frame_tasks[frame]
        corr = np.empty((n, self.num_pairs, 15, 15), np.float32)
        combo = np.empty((n, self.strip_height, 15), np.float32)
        disp = np.empty((n, _abi.DISP_LAYERS), np.float32)
        self._chk(self.lib.clt_get_corr2d_frame(self.h, frame, _ptr(corr), _ptr(combo), _ptr(disp)))
        return corr, combo, disp

    def getTexturesFrame(self, frame):
        n = self.frame_tasks[frame]
        planes = (4 + self.NUM_CAMS + 4) if self.keep_weights else 4
        tiles = np.empty((n, planes, 16, 16), np.float32)
        stats = np.empty((n, 4 * self.NUM_CAMS), np.float32)
        self._chk(self.lib.clt_get_textures_frame(self.h, frame, _ptr(tiles), _ptr(stats)))
        return tiles, stats

    def getRBGFrame(self, frame, ncam):
        out = np.empty((self.NUM_COLORS, (self.IMG_HEIGHT + 8) * (self.IMG_WIDTH + 8)), np.float32)
        self._chk(self.lib.clt_get_rbg_frame(self.h, frame, ncam, _ptr(out)))
        return out
