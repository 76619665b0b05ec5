"""CLT kernel file formats.

* The flat per-colour image of QuadCLT.flattenCLTKernels (QC:670-745) / extractCLTKernels
  (QC:790-857): each kernel tile is a 16 x 17 block, the four 8x8 quadrants (CC | SC over CS | SS)
  in the reference's CPU orientation and the 8 CltExtra values in the last row.  The conversion
  is the C-ABI's host code (clt_kernels_to_flat / clt_kernels_from_flat).
* The binary dumps of TwoQuadCLT.saveFloatKernels (TQ:1076-1177, little-endian) and
  saveFloatKernelsBigEndian (TQ:1181-1279): per camera ``<prefix>_chn<c>[_transposed].kernel``
  (float per ty, tx, colour, quadrant, 64 coefficients; "_transposed" = the [v][h] upload layout
  of setConvolutionKernels, otherwise the CPU orientation), ``.kernel_offsets`` (8 CltExtra floats
  per ty, tx, colour), ``<prefix>_chn<c>.bayer`` (the sum of the colour planes, one float per
  pixel) and ``<prefix>_chn<c>.portsxy`` (x, y per tile).

Kernels are passed as the clt_set_kernels layout: [KY][KX][K][4][8][8] float32 ([v][h] blocks)
with offsets [KY][KX][K][8].
"""
import os

import numpy as np

from . import _abi
from ._abi import CltError


def _ptr(a):
    return a.ctypes.data


def _chk(rc, what):
    if rc != 0:
        raise CltError(rc, what)


def flatten_clt_kernels(kernels, offsets):
    """QC:670-745: (kernels [KY][KX][K][4][8][8], offsets [KY][KX][K][8]) -> [K][17 KY][16 KX]."""
    k = np.ascontiguousarray(kernels, np.float32)
    o = np.ascontiguousarray(offsets, np.float32)
    if k.ndim != 6 or k.shape[3:] != (4, 8, 8) or o.shape != k.shape[:3] + (8,):
        raise ValueError("kernels must be [KY][KX][K][4][8][8] with offsets [KY][KX][K][8]")
    ky, kx, colors = k.shape[:3]
    flat = np.empty((colors, 17 * ky, 16 * kx), np.float32)
    _chk(_abi.load().clt_kernels_to_flat(_ptr(k), _ptr(o), kx, ky, colors, _ptr(flat)), "clt_kernels_to_flat")
    return flat


def extract_clt_kernels(flat):
    """QC:790-857: [K][17 KY][16 KX] flat images -> (kernels, offsets)."""
    f = np.ascontiguousarray(flat, np.float32)
    if f.ndim != 3 or f.shape[1] % 17 or f.shape[2] % 16:
        raise ValueError("flat kernels must be [K][17 KY][16 KX]")
    colors, ky, kx = f.shape[0], f.shape[1] // 17, f.shape[2] // 16
    k = np.empty((ky, kx, colors, 4, 8, 8), np.float32)
    o = np.empty((ky, kx, colors, 8), np.float32)
    _chk(_abi.load().clt_kernels_from_flat(_ptr(f), kx, ky, colors, _ptr(k), _ptr(o)), "clt_kernels_from_flat")
    return k, o


def _dtype(big_endian):
    return np.dtype(">f4" if big_endian else "<f4")


def kernel_paths(prefix, cam, transpose):
    """TQ:1083-1084 file names of one camera's kernel dump."""
    stem = "%s_chn%d%s" % (prefix, cam, "_transposed" if transpose else "")
    return stem + ".kernel", stem + ".kernel_offsets"


def save_float_kernels(prefix, kernels=None, offsets=None, image_data=None, port_xy=None, transpose=True,
                       big_endian=False):
    """TQ:1076-1177 (big_endian: TQ:1181-1279).  kernels / offsets: per camera arrays in the
    clt_set_kernels layout; image_data: per camera [K][H*W] colour planes (their sum is written) or
    [H][W] mosaics; port_xy: [tiles][cams][2]."""
    dt = _dtype(big_endian)
    if kernels is not None:
        for cam, (k, o) in enumerate(zip(kernels, offsets)):
            k = np.asarray(k, np.float32)
            if not transpose:   # CPU orientation: the transpose of each [v][h] block
                k = np.swapaxes(k, -1, -2)
            kp, op = kernel_paths(prefix, cam, transpose)
            np.ascontiguousarray(k, dt).tofile(kp)
            np.ascontiguousarray(o, dt).tofile(op)
    if image_data is not None:
        for cam, d in enumerate(image_data):
            d = np.asarray(d)
            if d.ndim == 2 and d.shape[0] in (1, 2, 3) and d.shape[1] > 3:
                d = d.astype(np.float64).sum(axis=0)   # TQ:1150 sums the colour planes in double
            np.ascontiguousarray(d, np.float64).astype(dt).tofile("%s_chn%d.bayer" % (prefix, cam))
    if port_xy is not None:
        pxy = np.asarray(port_xy, np.float64)
        for cam in range(pxy.shape[1]):
            np.ascontiguousarray(pxy[:, cam, :2]).astype(dt).tofile("%s_chn%d.portsxy" % (prefix, cam))


def read_float_kernels(prefix, cam, kx, ky, colors, transpose=True, big_endian=False):
    """Read one camera's kernel dump back into the clt_set_kernels layout."""
    dt = _dtype(big_endian)
    kp, op = kernel_paths(prefix, cam, transpose)
    nk = ky * kx * colors
    for path, n in ((kp, nk * 256), (op, nk * 8)):
        if os.path.getsize(path) != n * 4:
            raise ValueError("%s: %d bytes, expected %d for a %dx%dx%d kernel grid" %
                             (path, os.path.getsize(path), n * 4, kx, ky, colors))
    k = np.fromfile(kp, dt).astype(np.float32).reshape(ky, kx, colors, 4, 8, 8)
    if not transpose:
        k = np.ascontiguousarray(np.swapaxes(k, -1, -2))
    o = np.fromfile(op, dt).astype(np.float32).reshape(ky, kx, colors, 8)
    return k, o


def read_bayer(prefix, cam, width, height, big_endian=False):
    """One camera's ``.bayer`` dump as a [height][width] mosaic."""
    return np.fromfile("%s_chn%d.bayer" % (prefix, cam), _dtype(big_endian)).astype(np.float32).reshape(height, width)


def read_ports_xy(prefix, cam, big_endian=False):
    """One camera's ``.portsxy`` dump as [tiles][2]."""
    return np.fromfile("%s_chn%d.portsxy" % (prefix, cam), _dtype(big_endian)).astype(np.float32).reshape(-1, 2)
