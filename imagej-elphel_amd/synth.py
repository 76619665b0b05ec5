"""Synthetic inputs of the tile processor: textures, camera mosaics, CLT aberration kernels, tasks.

Everything is seeded (numpy ``default_rng``) so the GPU path and the CPU oracle see identical
float32 inputs.

* Scene texture: 1/f amplitude spectrum random field (DC removed) rescaled to 0..4095, one per
  colour with a shared component so colours correlate.
* Camera ``i`` image: the master texture resampled at ``q + d(q) * rXY_i`` (the object seen at
  master position ``p`` appears at ``p - d * rXY_i`` in camera ``i``), bilinear, Poisson noise.
* Bayer mosaic with the given red phase; green scaled by 0.5 as the reference does before the GPU
  (TwoQuadCLT.java:2062-2067, QuadCLT.java:180-183).
* Aberration kernels: per camera / kernel tile / colour Gaussian PSF (sigma 0.6..1.2 px) on the
  2x oversampled 64x64 grid, taken through the reference kernel chain
  clt_convert_double_kernel -> clt_normalize_kernel -> clt_symmetrize_kernel -> clt_dtt3_kernel
  -> clt_fill_coord_corr (ImageDtt.java:4625-4734, 4792-4851; QuadCLT.java:433-668), then laid out
  for upload (8x8 blocks transposed to [v][h], GPUTileProcessor.java:381-435).
"""
import numpy as np

from . import geometry


def texture_1f(h, w, rng, lo=0.0, hi=4095.0):
    """Random texture with a 1/f amplitude spectrum, DC removed, rescaled to [lo, hi]."""
    noise = rng.standard_normal((h, w))
    spec = np.fft.rfft2(noise)
    fy = np.fft.fftfreq(h)[:, None]
    fx = np.fft.rfftfreq(w)[None, :]
    f = np.sqrt(fx * fx + fy * fy)
    f[0, 0] = 1.0
    spec /= f
    spec[0, 0] = 0.0
    t = np.fft.irfft2(spec, s=(h, w))
    t = (t - t.min()) / (t.max() - t.min())
    return lo + (hi - lo) * t


def smooth_field(ny, nx, rng, lo, hi, terms=6):
    """Smooth random field on a ny x nx grid (sum of low-order random cosines) in [lo, hi]."""
    yy, xx = np.mgrid[0:ny, 0:nx]
    v = np.zeros((ny, nx))
    for _ in range(terms):
        kx, ky = rng.uniform(0.2, 2.5, 2) * np.pi / np.array([nx, ny])
        ph = rng.uniform(0, 2 * np.pi)
        v += rng.uniform(0.5, 1.0) * np.cos(kx * xx + ky * yy * rng.choice([-1, 1]) + ph)
    v = (v - v.min()) / max(v.max() - v.min(), 1e-12)
    return lo + (hi - lo) * v


def bayer_colors(h, w, red_row=0, red_col=1):
    """Colour index of every mosaic pixel: 0 red, 1 blue, 2 green (EyesisCorrections.java:2219)."""
    yy, xx = np.mgrid[0:h, 0:w]
    ry = (yy & 1) == red_row
    rx = (xx & 1) == red_col
    return np.where(ry & rx, 0, np.where(~ry & ~rx, 1, 2)).astype(np.int8)


def camera_images(width, height, rxy, disparity_tiles, rng, num_colors=3, red_row=0, red_col=1,
                  noise=True, margin=None, green_scale=0.5):
    """Mosaic images [C][H][W] float32 of one scene seen by C cameras.

    disparity_tiles: scalar or [H/8][W/8] per-tile disparity (pixels)."""
    from scipy.ndimage import map_coordinates, zoom
    rxy = np.asarray(rxy, np.float64)
    C = rxy.shape[0]
    d_t = np.broadcast_to(np.asarray(disparity_tiles, np.float64), (height // 8, width // 8))
    dmax = float(np.abs(d_t).max()) if d_t.size else 0.0
    m = margin if margin is not None else int(np.ceil(dmax * np.abs(rxy).max())) + 8
    H2, W2 = height + 2 * m, width + 2 * m
    base = texture_1f(H2, W2, rng)
    if num_colors == 3:
        gains = (0.85, 0.65, 1.0)   # R, B, G
        planes = [np.clip(g * base + 0.12 * (texture_1f(H2, W2, rng) - 2047.5), 0, 4095) for g in gains]
    else:
        planes = [base]
    # per-pixel disparity: bilinear upsample of the tile field (tile centres at 8*t+4)
    if np.all(d_t == d_t.flat[0]):
        dpix = np.full((height, width), d_t.flat[0])
    else:
        dpix = zoom(d_t, 8, order=1, mode="nearest")[:height, :width]
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    colors = bayer_colors(height, width, red_row, red_col) if num_colors == 3 else None
    out = np.empty((C, height, width), np.float32)
    for c in range(C):
        sy = yy + dpix * rxy[c, 1] + m
        sx = xx + dpix * rxy[c, 0] + m
        samp = [map_coordinates(p, [sy, sx], order=1, mode="nearest") for p in planes]
        if num_colors == 3:
            img = np.choose(colors, samp)
        else:
            img = samp[0]
        img = np.clip(img, 0, None)
        if noise:
            img = rng.poisson(img).astype(np.float64)
        if num_colors == 3 and green_scale != 1.0:
            img = np.where(colors == 2, img * green_scale, img)
        out[c] = img
    return out


# ---------------------------------------------------------------- CLT kernels
def _dct3e_mats():
    """DCT-IIIe / DST-IIIe matrices (DttRad2.java:970-991, 1032-1052)."""
    n = 8
    sc = np.sqrt(2.0 / n)
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    ec = np.where((j == 0) | (j == n - 1), 0.5, 1.0)
    es = np.where(j == n - 1, 0.5, 1.0)
    c3 = sc * ec * np.cos(j * (2 * k + 1) * np.pi / (2 * n))
    s3 = sc * es * np.sin((j + 1) * (2 * k + 1) * np.pi / (2 * n))
    return c3, s3


def clt_kernels(num_cams, num_colors, kernels_vert, kernels_hor, rng, sigma=(0.6, 1.2), jitter=0.3,
                kernel_step=16):
    """Kernels [C][KY][KX][K][4][8][8] ([v][h] blocks) and offsets [C][KY][KX][K][8], float32."""
    C, K, KY, KX = num_cams, num_colors, kernels_vert, kernels_hor
    n = C * KY * KX * K
    sig = rng.uniform(sigma[0], sigma[1], n)
    # smooth per-camera drift of the PSF centre over the field plus per-kernel jitter (pixels)
    drift = np.stack([smooth_field(KY, KX, rng, -0.4, 0.4) for _ in range(2 * C)]).reshape(C, 2, KY, KX)
    ox = (np.broadcast_to(drift[:, 0, :, :, None], (C, KY, KX, K)).reshape(n) + rng.uniform(-jitter, jitter, n))
    oy = (np.broadcast_to(drift[:, 1, :, :, None], (C, KY, KX, K)).reshape(n) + rng.uniform(-jitter, jitter, n))
    grid = (np.arange(64) - 32) * 0.5                       # oversampled grid, sensor pixels
    gx = np.exp(-((grid[None, :] - ox[:, None]) ** 2) / (2 * sig[:, None] ** 2))
    gy = np.exp(-((grid[None, :] - oy[:, None]) ** 2) / (2 * sig[:, None] ** 2))
    # clt_convert_double_kernel (ID:4625-4672): centroid, 15x15 downsample with 3x3 weights
    jj = np.arange(-32, 32)
    cxo = (gx * jj).sum(1) / gx.sum(1)
    cyo = (gy * jj).sum(1) / gy.sum(1)

    def src_center(c):
        s = np.floor(c + 0.5).astype(int) + 32
        s = np.where(s < 16, 15, np.where(s > 48, 48, s))
        return s

    sx, sy = src_center(cxo), src_center(cyo)
    off = 2 * np.arange(-7, 8)

    def down(g, s):
        idx = s[:, None] + off[None, :]
        r = np.arange(g.shape[0])[:, None]
        return 0.5 * g[r, idx - 1] + g[r, idx] + 0.5 * g[r, idx + 1]

    # The 15x15 kernel is separable, k15[y][x] = dy[y] * dx[x], so every step below is applied to
    # the two 15-vectors and the 8x8 quadrants are outer products.
    dx, dy = down(gx, sx), down(gy, sy)
    # clt_normalize_kernel (ID:4674-4696) with cos(pi i/16) cos(pi j/16) weights (QC:466-474)
    cw = np.cos(np.pi * np.abs(np.arange(-7, 8)) / 16)
    dy = dy * (16.0 / ((dy @ cw) * (dx @ cw)))[:, None]
    # clt_symmetrize_kernel (ID:4698-4723): with k0..k3 = k15[7 -+ i][7 -+ j],
    #   CC = (k0 + k1 + k2 + k3) / 4             = yp[i] xp[j] / 4
    #   SC = (-k0 + k1 - k2 + k3) / 4, j >= 1     = yp[i] xm[j] / 4
    #   CS = (-k0 - k1 + k2 + k3) / 4, i >= 1     = ym[i] xp[j] / 4
    #   SS = (-k0 + k1 - k2 + k3) / 4, i, j >= 1  = yp[i] xm[j] / 4 (the sign pattern of ID:4717)
    # the sine quadrants shifted by one (index 1..7 stored at 0..6, the last row/column zero)
    ii = np.arange(8)
    yp, ym = dy[:, 7 - ii] + dy[:, 7 + ii], dy[:, 7 + ii] - dy[:, 7 - ii]
    xp, xm = dx[:, 7 - ii] + dx[:, 7 + ii], dx[:, 7 + ii] - dx[:, 7 - ii]

    def shift1(v):
        return np.concatenate([v[:, 1:], np.zeros((v.shape[0], 1))], axis=1)

    vy = (yp, yp, shift1(ym), shift1(yp))
    vx = (xp, shift1(xm), xp, shift1(xm))
    # clt_dtt3_kernel (ID:4725-4734): dttt_iiie(quadrant q, mode q); upload layout [v][h] =
    # M_v . X . M_h^T with M_h = (q & 1 ? SIIIe : CIIIe), M_v = (q & 2 ? SIIIe : CIIIe); for
    # X = outer(vy, vx) / 4 that is outer(M_v vy, M_h vx) / 4
    c3, s3 = _dct3e_mats()
    kern = np.empty((n, 4, 8, 8))
    for q in range(4):
        mh = s3 if q & 1 else c3
        mv = s3 if q & 2 else c3
        kern[:, q] = 0.25 * (vy[q] @ mv.T)[:, :, None] * (vx[q] @ mh.T)[:, None, :]
    # extras: data_x, data_y, center_x, center_y (ID:4668-4671), derivatives (ID:4792-4851)
    extra = np.zeros((n, 8))
    extra[:, 0] = 0.5 * (sx - 32)
    extra[:, 1] = 0.5 * (sy - 32)
    extra[:, 2] = 0.5 * cxo
    extra[:, 3] = 0.5 * cyo
    extra = extra.reshape(C, KY, KX, K, 8)
    for c in range(C):
        for col in range(K):
            ctr_x = extra[c, :, :, col, 2]
            ctr_y = extra[c, :, :, col, 3]
            d = _fill_coord_corr(ctr_x, ctr_y, kernel_step)
            extra[c, :, :, col, 4:8] = d
    return (kern.reshape(C, KY, KX, K, 4, 8, 8).astype(np.float32), extra.astype(np.float32))


def raw_kernel_stack(num_colors, kernels_vert, kernels_hor, rng, ks=64, sigma=(0.7, 1.4)):
    """Double-resolution aberration kernels as calculateCLTKernel (QC:433-668) reads them: the
    stack [K][KY*ks][KX*ks] float32, kernel (tx, ty) of colour k in rows ty*ks.., columns tx*ks..
    Each PSF is a sum of two anisotropic rotated Gaussians (a core and a weaker coma-like lobe)
    with a field-dependent centre, so the kernels are neither separable nor symmetric."""
    K, KY, KX = num_colors, kernels_vert, kernels_hor
    n = K * KY * KX
    g = (np.arange(ks) - ks // 2) * 0.5   # sensor pixels
    drift = np.stack([smooth_field(KY, KX, rng, -0.8, 0.8) for _ in range(2)])
    cx = np.broadcast_to(drift[0][None], (K, KY, KX)).reshape(n) + rng.uniform(-0.3, 0.3, n)
    cy = np.broadcast_to(drift[1][None], (K, KY, KX)).reshape(n) + rng.uniform(-0.3, 0.3, n)
    out = np.zeros((n, ks, ks), np.float64)
    yy, xx = g[:, None], g[None, :]
    for lobe, amp in ((0, 1.0), (1, 0.35)):
        sa = rng.uniform(sigma[0], sigma[1], n) * (1.0 + lobe)
        sb = sa * rng.uniform(0.6, 1.0, n)
        th = rng.uniform(0, np.pi, n)
        ox = cx + lobe * rng.uniform(-1.0, 1.0, n)
        oy = cy + lobe * rng.uniform(-1.0, 1.0, n)
        for c0 in range(0, n, 512):
            c = slice(c0, min(n, c0 + 512))
            dx, dy = xx[None] - ox[c, None, None], yy[None] - oy[c, None, None]
            ct, st = np.cos(th[c])[:, None, None], np.sin(th[c])[:, None, None]
            u = (dx * ct + dy * st) / sa[c, None, None]
            v = (dy * ct - dx * st) / sb[c, None, None]
            out[c] += amp * np.exp(-0.5 * (u * u + v * v))
    return out.reshape(K, KY, KX, ks, ks).transpose(0, 1, 3, 2, 4).reshape(K, KY * ks, KX * ks).astype(np.float32)


def _fill_coord_corr(cx, cy, kstep):
    """clt_fill_coord_corr (ImageDtt.java:4792-4851): least-squares slopes of the kernel centres
    over the 3x3 neighbourhood -> dxc_dx, dxc_dy, dyc_dx, dyc_dy."""
    ny, nx = cx.shape
    acc = {k: np.zeros((ny, nx)) for k in ("s0", "sx", "sx2", "sy", "sy2", "sz", "sxz", "syz", "sw", "sxw", "syw")}
    for dty in (-1, 0, 1):
        for dtx in (-1, 0, 1):
            ys = slice(max(0, -dty), ny - max(0, dty))
            xs = slice(max(0, -dtx), nx - max(0, dtx))
            ysn = slice(max(0, dty), ny - max(0, -dty))
            xsn = slice(max(0, dtx), nx - max(0, -dtx))
            z = cx[ysn, xsn]
            w = cy[ysn, xsn]
            acc["s0"][ys, xs] += 1
            acc["sx"][ys, xs] += dtx
            acc["sx2"][ys, xs] += dtx * dtx
            acc["sy"][ys, xs] += dty
            acc["sy2"][ys, xs] += dty * dty
            acc["sz"][ys, xs] += z
            acc["sxz"][ys, xs] += dtx * z
            acc["syz"][ys, xs] += dty * z
            acc["sw"][ys, xs] += w
            acc["sxw"][ys, xs] += dtx * w
            acc["syw"][ys, xs] += dty * w
    a = acc
    denom_x = (a["sx2"] * a["s0"] - a["sx"] * a["sx"]) * kstep
    denom_y = (a["sy2"] * a["s0"] - a["sy"] * a["sy"]) * kstep
    return np.stack([(a["sxz"] * a["s0"] - a["sz"] * a["sx"]) / denom_x,
                     (a["syz"] * a["s0"] - a["sz"] * a["sy"]) / denom_y,
                     (a["sxw"] * a["s0"] - a["sw"] * a["sx"]) / denom_x,
                     (a["syw"] * a["s0"] - a["sw"] * a["sy"]) / denom_y], axis=-1)


# ---------------------------------------------------------------- whole scenes
class Scene:
    """All inputs of one frame: geometry, images, kernels, tasks and the ground truth."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def make_scene(width, height, num_cams=4, num_colors=3, rxy=None, disparity=3.0, seed=0,
               red_row=0, red_col=1, kernels=True, task_fraction=1.0, target_disparity=None,
               noise=True, flags=0xff):
    """One frame.  disparity: scalar, [TY][TX] field, or ('field', lo, hi) for a smooth random
    field; target_disparity: disparity used to build the tasks (default: the ground truth);
    kernels: True (random kernels of this camera set), False (none) or a (kernels, offsets) pair
    from clt_kernels (frames of one camera set share its kernels)."""
    rng = np.random.default_rng(seed)
    tiles_x, tiles_y = width // 8, height // 8
    if rxy is None:
        rxy = geometry.QUAD_RXY if num_cams == 4 else geometry.circle_rxy(num_cams)
    rxy = np.asarray(rxy, np.float64)
    if isinstance(disparity, tuple):
        _, lo, hi = disparity
        d_tiles = smooth_field(tiles_y, tiles_x, rng, lo, hi)
    else:
        d_tiles = np.broadcast_to(np.asarray(disparity, np.float64), (tiles_y, tiles_x)).copy()
    images = camera_images(width, height, rxy, d_tiles, rng, num_colors, red_row, red_col, noise=noise)
    kh, kv = width // 16 + 2, height // 16 + 2
    kern = offs = None
    if isinstance(kernels, tuple):
        kern, offs = kernels
    elif kernels:
        kern, offs = clt_kernels(num_cams, num_colors, kv, kh, rng)
    mask = None
    if task_fraction < 1.0:
        mask = rng.random((tiles_y, tiles_x)) < task_fraction
    td = d_tiles if target_disparity is None else np.broadcast_to(target_disparity, d_tiles.shape)
    tasks = geometry.full_frame_tasks(width, height, td, rxy, flags=flags, mask=mask)
    return Scene(width=width, height=height, num_cams=num_cams, num_colors=num_colors, rxy=rxy,
                 red_row=red_row, red_col=red_col, images=images, kernels=kern, offsets=offs,
                 tasks=tasks, disparity=d_tiles, kernels_hor=kh, kernels_vert=kv, seed=seed,
                 mask=mask)


class GeometryScene:
    """A quad scene rendered through the port geometry with given extrinsics (no aberrations)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def make_geometry_scene(width, height, extrinsic, seed=0, disparity=("field", 1.0, 4.0), focal_pix=4.5 / 0.0022,
                        red_row=0, red_col=1, noise=False):
    """Camera c pixel q shows the master texture at the point p with port_map(p, d(p))_c = q
    (solved by fixed-point steps); d(p) a smooth per-tile field, bilinear between tile centres."""
    from scipy.ndimage import map_coordinates, zoom
    rng = np.random.default_rng(seed)
    tiles_y, tiles_x = height // 8, width // 8
    if isinstance(disparity, tuple):
        d_t = smooth_field(tiles_y, tiles_x, rng, disparity[1], disparity[2])
    else:
        d_t = np.full((tiles_y, tiles_x), float(disparity))
    rxy = np.array(geometry.QUAD_RXY, np.float64)
    rots = geometry.corr_rot_matrices(extrinsic)
    m = 24
    H2, W2 = height + 2 * m, width + 2 * m
    base = texture_1f(H2, W2, rng)
    gains = (0.85, 0.65, 1.0)
    planes = [np.clip(g * base + 0.12 * (texture_1f(H2, W2, rng) - 2047.5), 0, 4095) for g in gains]
    dpix = zoom(d_t, 8, order=1, mode="nearest")[:height, :width]
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    colors = bayer_colors(height, width, red_row, red_col)
    images = np.empty((4, height, width), np.float32)
    for c in range(4):
        px, py = xx + dpix * rxy[c, 0], yy + dpix * rxy[c, 1]
        for _ in range(4):
            dp = map_coordinates(d_t, [(py - 4) / 8, (px - 4) / 8], order=1, mode="nearest")
            q = geometry.port_map(px, py, dp, rxy[c:c + 1], rots[c:c + 1], focal_pix, width, height)[..., 0, :]
            px, py = px + (xx - q[..., 0]), py + (yy - q[..., 1])
        samp = [map_coordinates(pl, [py + m, px + m], order=3, mode="nearest") for pl in planes]
        img = np.clip(np.choose(colors, samp), 0, None)
        if noise:
            img = rng.poisson(img).astype(np.float64)
        images[c] = np.where(colors == 2, img * 0.5, img)
    ty, tx = np.mgrid[0:tiles_y, 0:tiles_x]
    return GeometryScene(width=width, height=height, images=images, extrinsic=np.asarray(extrinsic, np.float64),
                         disparity_tiles=d_t, tx=tx.ravel(), ty=ty.ravel(), focal_pix=focal_pix, rots=rots)
