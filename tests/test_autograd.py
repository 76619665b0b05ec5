"""The adjoint wrapped in a torch.autograd.Function (verification harness only): torch's reverse
pass returns the hand-written gradients, and they agree with finite differences of the forward."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _functions(torch, tp, sc, tiles):
    class ExtrinsicDisparityLoss(torch.autograd.Function):
        """Summed correlation-peak loss of the tasks as a function of the 13 extrinsics and the
        task disparities; backward = the device adjoint (clt_exec_geometry_adjoint)."""

        @staticmethod
        def forward(ctx, extrinsic, disparity):
            geom = tp.defaultPortGeometry(sc.width, sc.height, extrinsic=extrinsic.detach().cpu().double().tolist())
            tp.setPortGeometry(geom)
            d = disparity.detach().cpu().numpy().astype(np.float32)
            tp.getTileSubcamOffsets(np.full(len(tiles), 0xFF, np.int32), sc.tx[tiles], sc.ty[tiles], d)
            ge, gd, loss = tp.getGeometryGradient()
            ctx.save_for_backward(torch.from_numpy(ge), torch.from_numpy(gd.astype(np.float64)))
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            ge, gd = ctx.saved_tensors
            return grad_out * ge, grad_out * gd

    return ExtrinsicDisparityLoss


def test_autograd_function_matches_adjoint_and_differences(pk):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(12)
    sc = pk.synth.make_geometry_scene(256, 192, rng.normal(0, 3e-4, 13), seed=12)
    inner = (sc.tx > 1) & (sc.tx < 30) & (sc.ty > 1) & (sc.ty < 22)
    tiles = np.nonzero(inner)[0]
    tp = pk.GPUTileProcessor(sc.width, sc.height, 4, 3, kernels=False)
    tp.setBayerImages(sc.images, force=True)
    tp.setCorrSigma(0.8)
    f = _functions(torch, tp, sc, tiles)
    e = torch.zeros(13, dtype=torch.float64, requires_grad=True)
    d = torch.tensor(sc.disparity_tiles.ravel()[tiles], dtype=torch.float64, requires_grad=True)
    loss = f.apply(e, d)
    loss.backward()
    ge, gd, l0 = tp.getGeometryGradient()
    assert np.array_equal(e.grad.numpy(), ge) and np.allclose(d.grad.numpy(), gd.astype(np.float64))
    # central differences of the forward (FP32 device path): coarse steps, loose tolerance
    for j, h in ((3, 2e-5), (6, 2e-4), (10, 2e-4)):
        ep, em = e.detach().clone(), e.detach().clone()
        ep[j] += h
        em[j] -= h
        fd = (f.apply(ep, d.detach()) - f.apply(em, d.detach())).item() / (2 * h)
        print("extrinsic %d: adjoint %.6g, differences %.6g" % (j, ge[j], fd))
        assert abs(fd - ge[j]) < 0.03 * max(abs(ge[j]), 1e-3 * np.abs(ge).max())
    tp.close()
