"""Port (sub-camera) tile centres and task records.

Host-side restatement of the task building of the reference GPU path:
``GPUTileProcessor.getTileSubcamOffsets`` (GPUTileProcessor.java:686-778) calling
``GeometryCorrection.getPortsCoordinatesAndDerivatives`` (GeometryCorrection.java:2159-2231) for
the radial-distortion-free case: un-distortion and re-distortion are identities, then for every
port ``i``

    v_i = R_i . (pXc - d * rX_i, pYc - d * rY_i, f)      (pXc, pYc relative to the image centre)
    p_i = (v_i.x, v_i.y) * f / v_i.z + pXY0_i

With identity rotations this is the closed form ``p_i = c - d * rXY_i + (pXY0_i - centre)``.
Vectorised over tiles with numpy (the reference runs one Java thread per core over tiles).
"""
import numpy as np

# GeometryCorrection.java:78 rXY of a quad camera (0 top-left, 1 top-right, 2 bottom-left,
# 3 bottom-right), in units of the disparity.
QUAD_RXY = np.array([[-0.5, -0.5], [0.5, -0.5], [-0.5, 0.5], [0.5, 0.5]])


def circle_rxy(n, radius=1.0):
    """Cameras on a circle: angle 2*pi*i/n (16-sensor arrays)."""
    a = 2 * np.pi * np.arange(n) / n
    return radius * np.stack([np.cos(a), np.sin(a)], axis=1)


def port_centers(width, height, tx, ty, disparity, rxy, pxy0=None, rots=None, focal_pix=None):
    """Per tile, per port pixel centres [n][C][2] (float64).

    tx, ty: tile indices [n]; disparity: [n] or scalar (pixels); rxy: [C][2];
    pxy0: [C][2] port image centres (default: the image centre); rots: [C][3][3] or None;
    focal_pix: focal length in pixels (needed with rotations)."""
    tx = np.asarray(tx, np.float64)
    ty = np.asarray(ty, np.float64)
    d = np.broadcast_to(np.asarray(disparity, np.float64), tx.shape)
    cx = tx * 8 + 4 - 0.5 * width
    cy = ty * 8 + 4 - 0.5 * height
    rxy = np.asarray(rxy, np.float64)
    C = rxy.shape[0]
    if pxy0 is None:
        pxy0 = np.tile([0.5 * width, 0.5 * height], (C, 1))
    px = cx[:, None] - d[:, None] * rxy[None, :, 0]
    py = cy[:, None] - d[:, None] * rxy[None, :, 1]
    if rots is not None:
        f = float(focal_pix)
        v = np.stack([px, py, np.full_like(px, f)], axis=-1)          # [n][C][3]
        rv = np.einsum("cij,ncj->nci", np.asarray(rots, np.float64), v)
        nz = f / rv[..., 2]
        px = rv[..., 0] * nz
        py = rv[..., 1] * nz
    out = np.empty(px.shape + (2,))
    out[..., 0] = px + pxy0[None, :, 0]
    out[..., 1] = py + pxy0[None, :, 1]
    return out


def make_tasks(tx, ty, centers, flags=0xff):
    """Pack TpTask records (GPUTileProcessor.java:132-166): int task, int txy, float xy[C][2]."""
    tx = np.asarray(tx, np.int32)
    ty = np.asarray(ty, np.int32)
    centers = np.asarray(centers, np.float32)
    n, C = centers.shape[0], centers.shape[1]
    rec = np.zeros((n, 2 + 2 * C), np.int32)
    rec[:, 0] = np.broadcast_to(np.asarray(flags, np.int32), (n,))
    rec[:, 1] = tx | (ty << 16)
    rec[:, 2:] = centers.reshape(n, 2 * C).view(np.int32)
    return rec


def full_frame_tasks(width, height, disparity, rxy, flags=0xff, mask=None, **kw):
    """setFullFrameImages (GPUTileProcessor.java:472-527): one task per tile, optional mask
    [tilesY][tilesX] selecting a sparse subset; disparity scalar or [tilesY][tilesX]."""
    tiles_x, tiles_y = width // 8, height // 8
    ty, tx = np.mgrid[0:tiles_y, 0:tiles_x]
    d = np.broadcast_to(np.asarray(disparity, np.float64), (tiles_y, tiles_x))
    sel = np.ones((tiles_y, tiles_x), bool) if mask is None else np.asarray(mask, bool)
    tx, ty, d = tx[sel], ty[sel], d[sel]
    centers = port_centers(width, height, tx, ty, d, rxy, **kw)
    return make_tasks(tx, ty, centers, flags)


def corr_rot_matrices(extrinsic, roll_deg=(0.0, 0.0, 0.0, 0.0)):
    """Per sensor rotation/zoom matrices [4][3][3] of a CorrVector (GC:1043-1076): tilt0..2,
    azimuth0..2, roll0..3, zoom0..2; sensor 3 tilt, azimuth and zoom are minus the sum of the
    others; factory rolls in degrees are added."""
    v = np.asarray(extrinsic, np.float64)
    out = np.zeros((4, 3, 3))
    for i in range(4):
        tl = v[i] if i < 3 else -(v[0] + v[1] + v[2])
        az = v[3 + i] if i < 3 else -(v[3] + v[4] + v[5])
        rl = v[6 + i] + np.deg2rad(roll_deg[i])
        zm = 1.0 + (v[10 + i] if i < 3 else -(v[10] + v[11] + v[12]))
        ca, sa, ct, st, cr, sr = np.cos(az), np.sin(az), np.cos(tl), np.sin(tl), np.cos(rl), np.sin(rl)
        m_az = np.array([[ca, 0, -sa], [0, 1, 0], [sa, 0, ca]])
        m_t = np.array([[1, 0, 0], [0, ct, st], [0, -st, ct]])
        m_r = np.array([[cr * zm, sr * zm, 0], [-sr * zm, cr * zm, 0], [0, 0, 1]])
        out[i] = m_r @ (m_t @ m_az)
    return out


def port_map(px, py, disparity, rxy, rots, focal_pix, width, height, pxy0=None):
    """Undistorted port model at arbitrary points: pixel (px, py) of the virtual centre camera at
    disparity d seen by sensor c at [..][C][2] (GC:2181-2231 without radial distortion)."""
    px = np.asarray(px, np.float64)
    d = np.broadcast_to(np.asarray(disparity, np.float64), px.shape)
    rxy = np.asarray(rxy, np.float64)
    C = rxy.shape[0]
    if pxy0 is None:
        pxy0 = np.tile([0.5 * width, 0.5 * height], (C, 1))
    x = (px - 0.5 * width)[..., None] - d[..., None] * rxy[:, 0]
    y = (np.asarray(py, np.float64) - 0.5 * height)[..., None] - d[..., None] * rxy[:, 1]
    f = float(focal_pix)
    r = np.asarray(rots, np.float64)
    rx = r[:, 0, 0] * x + r[:, 0, 1] * y + r[:, 0, 2] * f
    ry = r[:, 1, 0] * x + r[:, 1, 1] * y + r[:, 1, 2] * f
    rz = r[:, 2, 0] * x + r[:, 2, 1] * y + r[:, 2, 2] * f
    out = np.empty(x.shape + (2,))
    out[..., 0] = rx * f / rz + pxy0[:, 0]
    out[..., 1] = ry * f / rz + pxy0[:, 1]
    return out
