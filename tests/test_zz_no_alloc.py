"""clt_exec_geometry_adjoint_noalloc (no device allocation, static scratch) against
clt_exec_geometry_adjoint on the same handle: the same kernels in the same order, so the
gradients and the loss are bit-identical; two calls in a row, the second on fewer tasks (the
scratch keeps the first call's contents), then the allocating entry point again; and the size
limit of the scratch."""
import numpy as np
import pytest

from helpers import log

pytestmark = pytest.mark.gpu


def _processor(pk, sc, tiles, ext):
    tp = pk.GPUTileProcessor(sc.width, sc.height, 4, 3)
    tp.setConvolutionKernels(sc.kernels, offsets=sc.offsets, force=True)
    tp.setBayerImages(sc.images, force=True)
    tp.setCorrSigma(0.8)
    tp.setPortGeometry(pk.GPUTileProcessor.defaultPortGeometry(sc.width, sc.height, extrinsic=list(ext)))
    tx, ty = tiles % (sc.width // 8), tiles // (sc.width // 8)
    tp.getTileSubcamOffsets(np.full(len(tiles), 0xFF, np.int32), tx, ty, sc.disparity[ty, tx].astype(np.float32))
    return tp


def _same(a, b):
    return all(np.array_equal(np.asarray(x).view(np.int64 if np.asarray(x).dtype == np.float64 else np.int32),
                              np.asarray(y).view(np.int64 if np.asarray(y).dtype == np.float64 else np.int32))
               for x, y in zip(a, b))


def test_noalloc_matches_allocating_entry_point(pk):
    sc = pk.synth.make_scene(256, 192, seed=31, disparity=("field", 0.0, 6.0))
    ext = np.random.default_rng(3).normal(0, 1e-3, 13)
    ntiles = (256 // 8) * (192 // 8)
    big = np.arange(ntiles)
    small = np.random.default_rng(4).choice(ntiles, 150, replace=False)
    tp = _processor(pk, sc, big, ext)
    for step, tiles in enumerate((big, small)):
        if step:
            tx, ty = tiles % 32, tiles // 32
            tp.getTileSubcamOffsets(np.full(len(tiles), 0xFF, np.int32), tx, ty, sc.disparity[ty, tx].astype(np.float32))
        log("case: no-allocation geometry adjoint, call %d, %d tasks" % (step + 1, len(tiles)))
        na = tp.getGeometryGradientNoAlloc()
        ref = tp.getGeometryGradient()
        na2 = tp.getGeometryGradientNoAlloc()
        log("  loss %.6f / %.6f, max |dl/de| diff %.3g, max |dl/dd| diff %.3g" % (
            na[2], ref[2], np.abs(na[0] - ref[0]).max(), np.abs(na[1] - ref[1]).max()))
        assert len(na[1]) == len(tiles)
        assert _same((na[0], na[1], np.float64(na[2])), (ref[0], ref[1], np.float64(ref[2])))
        assert _same((na2[0], na2[1], np.float64(na2[2])), (ref[0], ref[1], np.float64(ref[2])))
    tp.close()


def test_noalloc_rejects_more_tasks_than_the_scratch(pk):
    w, h = 2608, 2000                      # 326 x 250 = 81,500 tiles > 80,000
    tp = pk.GPUTileProcessor(w, h, 4, 3, kernels=False)
    tp.setBayerImages(np.zeros((4, h, w), np.float32), force=True)
    tp.setPortGeometry(pk.GPUTileProcessor.defaultPortGeometry(w, h))
    n = (w // 8) * (h // 8)
    assert n > tp.MAX_TASKS_NOALLOC
    t = np.arange(n)
    tp.getTileSubcamOffsets(np.full(n, 0xFF, np.int32), t % (w // 8), t // (w // 8), np.zeros(n, np.float32))
    with pytest.raises(pk.CltError, match="static scratch"):
        tp.getGeometryGradientNoAlloc()
    tp.close()
