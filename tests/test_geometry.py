"""Port geometry (GeometryCorrection.getPortsCoordinatesAndDerivatives, GC:2159-2314): the oracle
pinned by closed forms and finite differences on the CPU, the device kernel against the oracle."""
import ctypes

import numpy as np
import pytest

W, H = 2592, 1936
RNG_EXTR = [1e-3, -2e-3, 5e-4, 2e-3, 1e-3, -1e-3, 3e-3, -1e-3, 2e-3, 1e-3, 1e-3, -2e-3, 1e-3]
DIST = [0.02, -0.01, 0.005, 0.002, 0.0, 0.0, 0.0]


def _tiles(n, seed):
    rng = np.random.default_rng(seed)
    tx = rng.integers(0, W // 8, n)
    ty = rng.integers(0, H // 8, n)
    return tx, ty, rng.uniform(0.0, 20.0, n)


def test_zero_distortion_identity_closed_form(orc):
    """pXY = (c - W/2) - d * rXY + pXY0 when nothing is rotated or distorted (GC:2181-2231)."""
    pg = orc.default_port_geometry(W, H)
    tx, ty, d = _tiles(50, 1)
    px, py = tx * 8 + 4.0, ty * 8 + 4.0
    xy = orc.port_coords(pg, px, py, d)
    rxy = np.array(orc.QUAD_PORT_OFFSETS)
    exp = np.stack([px, py], -1)[:, None, :] - d[:, None, None] * rxy[None]
    assert np.abs(xy - exp).max() < 1e-9


def test_single_axis_rotations_closed_forms(orc):
    f = 4.5 / 0.0022                         # focal length in pixels (GC:2189)
    x, y, d = 300.0, -200.0, 0.0
    px, py = W / 2 + x, H / 2 + y
    th = 0.01
    # roll of sensor 0: (x, y) -> (c x + s y, -s x + c y)
    xy = orc.port_coords(orc.default_port_geometry(W, H, extrinsic=[0] * 6 + [th, 0, 0, 0, 0, 0, 0]), px, py, d)[0]
    c, s = np.cos(th), np.sin(th)
    assert np.allclose(xy[0] - [W / 2, H / 2], [c * x + s * y, -s * x + c * y], atol=1e-9)
    assert np.allclose(xy[1] - [W / 2, H / 2], [x, y], atol=1e-9)
    # zoom of sensor 1 scales, sensor 3 gets minus the sum of zooms 0..2
    xy = orc.port_coords(orc.default_port_geometry(W, H, extrinsic=[0] * 11 + [th, 0]), px, py, d)[0]
    assert np.allclose(xy[1] - [W / 2, H / 2], [(1 + th) * x, (1 + th) * y], atol=1e-9)
    assert np.allclose(xy[3] - [W / 2, H / 2], [(1 - th) * x, (1 - th) * y], atol=1e-9)
    # azimuth of sensor 2 turns the view about the vertical axis: x'' = f tan(atan(x / f) - az)
    xy = orc.port_coords(orc.default_port_geometry(W, H, extrinsic=[0] * 5 + [th] + [0] * 7), px, py, d)[0]
    zc = np.sin(th) * x + np.cos(th) * f
    assert np.allclose(xy[2] - [W / 2, H / 2], [f * np.tan(np.arctan(x / f) - th), f * y / zc], atol=1e-7)
    # tilt of sensor 0 about the horizontal axis: y'' = f tan(atan(y / f) + tilt)
    xy = orc.port_coords(orc.default_port_geometry(W, H, extrinsic=[th] + [0] * 12), px, py, d)[0]
    zc = -np.sin(th) * y + np.cos(th) * f
    assert np.allclose(xy[0] - [W / 2, H / 2], [f * x / zc, f * np.tan(np.arctan(y / f) + th)], atol=1e-7)


def test_rotation_matrices_are_zoomed_rotations(orc):
    pg = orc.default_port_geometry(W, H, extrinsic=RNG_EXTR, roll_deg=[0.5, -0.3, 0.2, 0.1])
    r, _ = orc.rot_matrices(pg)
    zooms = [1 + RNG_EXTR[10], 1 + RNG_EXTR[11], 1 + RNG_EXTR[12], 1 - sum(RNG_EXTR[10:13])]
    for i in range(4):
        q = np.diag([1 / zooms[i], 1 / zooms[i], 1.0]) @ r[i]
        assert np.abs(q @ q.T - np.eye(3)).max() < 1e-12


def test_reverse_distortion_round_trip(orc):
    """getRByRDist inverts the distortion polynomial (GC:2707-2830): r_nd = rD * table(rD)."""
    pg = orc.default_port_geometry(W, H, distortion=DIST)
    c = DIST
    for r in (0.05, 0.3, 0.7, 1.0, 1.4):
        k = 1 + sum(c[j] * (r ** (j + 1) - 1) for j in range(7))   # Rdist / R (GC:2217-2222)
        rd = r * k
        assert abs(rd * orc.r_by_rdist(pg, rd) - r) < 1e-6
    # undistort the tile centre, then distort it again: the port centre at zero disparity
    px, py = np.array([100.0, 1300.0, 2500.0]), np.array([50.0, 900.0, 1900.0])
    xy = orc.port_coords(pg, px, py, 0.0)
    assert np.abs(xy[:, 0] - np.stack([px, py], -1)).max() < 1e-3


def test_derivatives_match_central_differences(orc):
    kw = dict(distortion=DIST, roll_deg=[0.5, -0.3, 0.2, 0.1])
    pg = orc.default_port_geometry(W, H, extrinsic=RNG_EXTR, **kw)
    tx, ty, d = _tiles(20, 2)
    px, py = tx * 8 + 4.0, ty * 8 + 4.0
    _, der = orc.port_coords(pg, px, py, d, deriv=True)
    h = 1e-6
    for k in range(14):
        vals = []
        for sg in (1, -1):
            e = list(RNG_EXTR)
            dd = d.copy()
            if k < 13:
                e[k] += sg * h
            else:
                dd += sg * h
            vals.append(orc.port_coords(orc.default_port_geometry(W, H, extrinsic=e, **kw), px, py, dd))
        fd = (vals[0] - vals[1]).reshape(len(px), 8) / (2 * h)
        assert np.abs(fd - der[:, :, k]).max() < 1e-5 * max(1.0, np.abs(der[:, :, k]).max()), k


def test_abi_struct_and_defaults(pk, orc):
    assert ctypes.sizeof(pk._abi.clt_port_geometry) == ctypes.sizeof(orc.PortGeometry)
    p = pk.GPUTileProcessor.defaultPortGeometry(W, H)
    q = orc.default_port_geometry(W, H)
    assert bytes(p) == bytes(q)


@pytest.mark.gpu
def test_device_port_centres_match_oracle(pk, orc):
    tp = pk.GPUTileProcessor(W, H, 4, 3, kernels=False)
    kw = dict(distortion=DIST, extrinsic=RNG_EXTR, roll_deg=[0.5, -0.3, 0.2, 0.1])
    tp.setPortGeometry(pk.GPUTileProcessor.defaultPortGeometry(W, H, **kw))
    tx, ty, d = _tiles(5000, 3)
    d = d.astype(np.float32)
    word = np.full(len(tx), 0xFF, np.int32)
    word[::7] = 0
    tp.getTileSubcamOffsets(word, tx, ty, d)
    rec = tp.getTasks()
    g_xy = rec[:, 2:].view(np.float32).reshape(-1, 4, 2)
    o_xy, o_der = orc.port_coords(orc.default_port_geometry(W, H, **kw), tx * 8 + 4.0, ty * 8 + 4.0,
                                  d.astype(np.float64), deriv=True)
    live = word != 0
    err = np.abs(g_xy[live] - o_xy[live].astype(np.float32)).max()
    print("port centres: max |gpu - oracle| = %.3g px over %d tasks" % (err, live.sum()))
    assert err <= 1e-3
    assert np.all(g_xy[~live] == 0) and np.array_equal(rec[:, 0], word)
    g_der = tp.getPortDerivatives()
    rel = np.abs(g_der[live] - o_der[live]).max() / np.abs(o_der[live]).max()
    print("port derivatives: max relative error %.3g" % rel)
    assert rel < 1e-9
    tp.close()


@pytest.mark.gpu
@pytest.mark.parametrize("per_tile", [False, True])
def test_set_full_frame_images(pk, orc, per_tile):
    """setFullFrameImages (GTP:472-527): every tile of the frame in row-major order, at one
    disparity or one per tile, centres on the device equal to the oracle's port coordinates."""
    w, h = 320, 240
    tiles_x, tiles_y = w // 8, h // 8
    n = tiles_x * tiles_y
    d = np.linspace(0.0, 12.0, n).astype(np.float32) if per_tile else 2.5
    tp = pk.GPUTileProcessor(w, h, 4, 3, kernels=False)
    kw = dict(distortion=DIST, extrinsic=RNG_EXTR)
    tp.setPortGeometry(pk.GPUTileProcessor.defaultPortGeometry(w, h, **kw))
    rec = tp.setFullFrameImages(d)
    tp.close()
    assert rec.shape == (n, 10) and (rec[:, 0] == 0xFF).all()
    tx, ty = rec[:, 1] & 0xFFFF, rec[:, 1] >> 16
    assert np.array_equal(ty * tiles_x + tx, np.arange(n))
    o_xy = orc.port_coords(orc.default_port_geometry(w, h, **kw), tx * 8 + 4.0, ty * 8 + 4.0,
                           np.broadcast_to(np.asarray(d, np.float64), (n,)))
    g_xy = rec[:, 2:].view(np.float32).reshape(-1, 4, 2)
    err = np.abs(g_xy - o_xy.astype(np.float32)).max()
    print("full frame port centres: max |gpu - oracle| = %.3g px over %d tiles" % (err, n))
    assert err <= 1e-3
