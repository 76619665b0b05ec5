// CLT kernel preparation on the device (QuadCLT.calculateCLTKernel, QC:433-668): the per-camera,
// per-colour grid of double-resolution aberration kernels (PSFs, 64x64 each) becomes the 4-quadrant
// CLT kernels and CltExtra records that convert_direct multiplies by.  One wavefront per kernel
// tile, FP64 like the reference: centroid, 15x15 downsample around it (ID:4625-4672), normalisation
// (ID:4674-4696), symmetrisation into the four quadrants (ID:4698-4723), DCT-IIIe/DST-IIIe of each
// (ID:4725-4734), the sensor offset (offsetKernelSensor ID:4781-4791); then the centre slopes over
// the kernel grid (clt_fill_coord_corr, ID:4792-4851) in a second kernel.
#pragma once

namespace clt {

// tab: CIIIe[k][j] (64), SIIIe[k][j] (64), normalisation window cos(pi i/16) cos(pi j/16) (64)
struct KPrepArgs {
    int ks;          // raw kernel size (64)
    int kx, ky;      // kernel grid of the camera
    int colors;
    int row0, rows;  // kernel rows held by the raw band
    int norm_kern;
};

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

// raw band: [colors][rows * ks][kx * ks]; kern: the camera's device kernels [ky][kx][colors][4][h][v];
// ctr: [colors][ky][kx][2] centres (double, for the slopes); offs: [ky][kx][colors][8]
__global__ __launch_bounds__(256) void k_prepare_kernels(const float* __restrict__ raw, const double* __restrict__ tab,
                                                         const float* __restrict__ sensor_xy, float* __restrict__ kern,
                                                         float* __restrict__ offs, double* __restrict__ ctr, KPrepArgs A) {
    __shared__ double s_k[4][225];
    __shared__ double s_q[4][4][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + wave;
    const int per_col = A.rows * A.kx;
    if (w >= per_col * A.colors) return;   // wave-uniform
    const int col = w / per_col, r = w - col * per_col;
    const int tyl = r / A.kx, tx = r - tyl * A.kx, ty = A.row0 + tyl;
    const int ks = A.ks, c0 = ks / 2;
    const size_t W = (size_t)A.kx * ks;
    const float* src = raw + (size_t)col * A.rows * ks * W + (size_t)tyl * ks * W + (size_t)tx * ks;   // extractOneKernel
    // centroid (ID:4633-4645)
    double sx = 0.0, sy = 0.0, s = 0.0;
    for (int i = 0; i < ks; i++)
        for (int j = lane; j < ks; j += 64) {
            const double d = src[(size_t)i * W + j];
            sx += (j - c0) * d;
            sy += (i - c0) * d;
            s += d;
        }
    sx = wave_sum_d(sx);
    sy = wave_sum_d(sy);
    s = wave_sum_d(s);
    int src_x = (int)floor(sx / s + 0.5) + c0, src_y = (int)floor(sy / s + 0.5) + c0;   // Math.round
    if (src_x < 16) src_x = 15;
    else if (src_x > ks - 16) src_x = ks - 16;
    if (src_y < 16) src_y = 15;
    else if (src_y > ks - 16) src_y = ks - 16;
    // 15x15 downsample with the 3x3 weights, in the reference's tap order
    double* k15 = s_k[wave];
    const int tap_r[9] = {0, -1, 0, 0, 1, -1, -1, 1, 1}, tap_c[9] = {0, 0, -1, 1, 0, -1, 1, -1, 1};
    const double tap_w[9] = {1.0, 0.5, 0.5, 0.5, 0.5, 0.25, 0.25, 0.25, 0.25};
    double part = 0.0;
    for (int o = lane; o < 225; o += 64) {
        const int i = o / 15 - 7, j = o % 15 - 7;
        const int rr = src_y + 2 * i, cc = src_x + 2 * j;
        double d = 0.0;
#pragma unroll
        for (int t = 0; t < 9; t++) d += tap_w[t] * (double)src[(size_t)(rr + tap_r[t]) * W + cc + tap_c[t]];
        k15[o] = d;
        part += d * tab[128 + abs(i) * 8 + abs(j)];
    }
    if (A.norm_kern) {   // ID:4674-4696
        const double norm = 16 * (1.0 / wave_sum_d(part));
        for (int o = lane; o < 225; o += 64) k15[o] *= norm;
    }
    wave_sync();
    // symmetrise (ID:4698-4723): lane = i * 8 + j
    {
        const int i = lane >> 3, j = lane & 7;
        const int ctr0 = 7 * 15 + 7;
        const double a = k15[ctr0 - i * 15 - j], b = k15[ctr0 - i * 15 + j], c = k15[ctr0 + i * 15 - j], d = k15[ctr0 + i * 15 + j];
        double* q = &s_q[wave][0][0];
        q[lane] = 0.25 * (a + b + c + d);
        if (j == 7) q[64 + i * 8 + 7] = 0.0;
        if (i == 7) q[128 + 7 * 8 + j] = 0.0;
        if (j == 7) q[192 + i * 8 + 7] = 0.0;
        if (i == 7) q[192 + 7 * 8 + j] = 0.0;
        wave_sync();
        if (j > 0) q[64 + i * 8 + j - 1] = 0.25 * (-a + b - c + d);
        if (i > 0) q[128 + (i - 1) * 8 + j] = 0.25 * (-a - b + c + d);
        if (i > 0 && j > 0) q[192 + (i - 1) * 8 + (j - 1)] = 0.25 * (-a + b - c + d);   // the reference's signs
    }
    wave_sync();
    // DCT-IIIe / DST-IIIe of quadrant q in mode q (ID:4725-4734, DR:648-666): pass 1 along the
    // columns (table by mode bit 0), pass 2 along the rows (mode bit 1); result [h][v]
    float* ko = kern + (((size_t)ty * A.kx + tx) * A.colors + col) * 256;
    for (int q = 0; q < 4; q++) {
        const double* x = s_q[wave][q];
        const double* m0 = tab + ((q & 1) ? 64 : 0);
        const double* m1 = tab + ((q & 2) ? 64 : 0);
        const int a = lane >> 3, i = lane & 7;   // pass 1: t[a][i] = sum_j m0[a][j] x[i][j]
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < 8; j++) t += m0[a * 8 + j] * x[i * 8 + j];
        wave_sync();
        double* tt = k15;   // reuse: the 15x15 is no longer needed
        tt[lane] = t;
        wave_sync();
        const int b = lane & 7;   // pass 2: y[a][b] = sum_i m1[b][i] t[a][i]
        double y = 0.0;
#pragma unroll
        for (int ii = 0; ii < 8; ii++) y += m1[b * 8 + ii] * tt[a * 8 + ii];
        ko[q * 64 + lane] = (float)y;
        wave_sync();
    }
    if (lane == 0) {   // CltExtra items 0..3 (+ the sensor offset); 4..7 by k_fill_coord_corr
        double dx = 0.0, dy = 0.0;
        if (sensor_xy) {
            dx = -(double)sensor_xy[((size_t)ty * A.kx + tx) * 2];
            dy = -(double)sensor_xy[((size_t)ty * A.kx + tx) * 2 + 1];
        }
        const double data_x = 0.5 * (src_x - c0) + dx, data_y = 0.5 * (src_y - c0) + dy;
        const double cx = 0.5 * (sx / s) + dx, cy = 0.5 * (sy / s) + dy;
        float* e = offs + (((size_t)ty * A.kx + tx) * A.colors + col) * 8;
        e[0] = (float)data_x;
        e[1] = (float)data_y;
        e[2] = (float)cx;
        e[3] = (float)cy;
        double* cp = ctr + (((size_t)col * A.ky + ty) * A.kx + tx) * 2;
        cp[0] = cx;
        cp[1] = cy;
    }
}

// clt_fill_coord_corr (ID:4792-4851): one thread per (colour, kernel tile)
__global__ __launch_bounds__(256) void k_fill_coord_corr(const double* __restrict__ ctr, float* __restrict__ offs, int kx,
                                                         int ky, int colors, int kern_step) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= colors * ky * kx) return;
    const int col = t / (ky * kx), r = t - col * ky * kx, ty = r / kx, tx = r - ty * kx;
    double s0 = 0, sx = 0, sx2 = 0, sy = 0, sy2 = 0, sz = 0, sxz = 0, syz = 0, sw = 0, sxw = 0, syw = 0;
    for (int dty = -1; dty < 2; dty++) {
        const int y = ty + dty;
        if (y < 0 || y >= ky) continue;
        for (int dtx = -1; dtx < 2; dtx++) {
            const int x = tx + dtx;
            if (x < 0 || x >= kx) continue;
            const double* c = ctr + (((size_t)col * ky + y) * kx + x) * 2;
            s0 += 1;
            sx += dtx;
            sx2 += dtx * dtx;
            sy += dty;
            sy2 += dty * dty;
            sz += c[0];
            sxz += dtx * c[0];
            syz += dty * c[0];
            sw += c[1];
            sxw += dtx * c[1];
            syw += dty * c[1];
        }
    }
    const double denom_x = (sx2 * s0 - sx * sx) * kern_step, denom_y = (sy2 * s0 - sy * sy) * kern_step;
    float* e = offs + (((size_t)ty * kx + tx) * colors + col) * 8;
    e[4] = (float)((sxz * s0 - sz * sx) / denom_x);
    e[5] = (float)((syz * s0 - sz * sy) / denom_y);
    e[6] = (float)((sxw * s0 - sw * sx) / denom_x);
    e[7] = (float)((syw * s0 - sw * sy) / denom_y);
}

}  // namespace clt
