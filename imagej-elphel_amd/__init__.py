"""MI355X-native CLT tile processor (imagej-elphel's convert_direct / correlate2D / textures /
imclt_rbg path) behind the reference's GPUTileProcessor interface.

Import with ``importlib.import_module("imagej-elphel_amd")`` (the directory name has a hyphen).
The compute path is ``libclt_hip.so`` (hand-written gfx950 HIP kernels); there is no CPU fallback.
"""
from . import geometry, kernel_files, synth  # noqa: F401
from ._abi import CltError, load  # noqa: F401
from .tile_processor import GPUTileBatch, GPUTileProcessor, TpTask  # noqa: F401

__all__ = ["GPUTileProcessor", "GPUTileBatch", "TpTask", "CltError", "load", "geometry", "kernel_files", "synth"]
