// Two-quad rig (ImageDtt.clt_bi_quad, ID:8731-9504) on the device: a main quad camera (cameras
// 0..3) and an auxiliary one (4..7) converted together as one 8-camera frame.  Each quad runs the
// quad correlate2D kernel on its own cameras (which also writes the mean CLT of its 4 cameras,
// cltMixCameras C2:425-447) and the per-tile LMA in its tileCorrs form (ID:7581-7740: the fit over
// all pairs, the four one-direction fits, no foreground correction).  The inter-camera stage
// (correlateInterCamerasFD C2:401-448, tileInterCamCorrs ID:7370-7560) correlates the two means as
// one horizontal pair: k_corr_inter (correlation, strip, integer maximum, centre of mass or its
// notch form, single2dCM) and k_corr_lma_inter (the LMA and the 2-D polynomial dx, dy).  The
// low-texture pass (ID:9404-9480) averages the inter correlations of neighbouring tiles and runs
// the inter stage again; k_rig_bimap assembles the BIDISPARITY record (ID:89-114).
#pragma once

namespace clt {

constexpr int RIG_REC = 8;   // k_corr_inter -> k_corr_lma_inter record

struct RigArgs {
    float min_corr;      // min_corr, halved when no_int_x0 (ID:7404-7406)
    int cm_at_zero;      // centre of mass around 0 (no_int_x0, or the notch form) instead of the integer maximum
    int cm_rows;         // strip rows the centre of mass reads
    float cm_wy[16];     // per row: window_y * (row 0: 1, else 2) * wy_scale (C2:942-949, 994)
    float wndx[9];
    float cm_max_norm;   // cm_max_normalization (IDP:112)
    int iradius, ir2, square;   // single2dCM window (C2:1551-1553)
};

// One wavefront per task.  FROM_MIX: correlate the two quads' mean CLT (mix_main, mix_aux
// [T][3][4][64], one bin per lane) into corr_out [T][225]; else read the correlation from src
// (the low-texture averages).  Then the analysis of tileInterCamCorrs up to the LMA into rec_out.
template <bool FZ, bool FROM_MIX>
__global__ __launch_bounds__(256) void k_corr_inter(const float* __restrict__ mix_main, const float* __restrict__ mix_aux,
                                                    const float* __restrict__ src, const int32_t* __restrict__ tasks,
                                                    int task_words, int num_tasks, const float* __restrict__ lpf,
                                                    const CorrArgs* __restrict__ args_p, const RigArgs* __restrict__ rig_p,
                                                    float* __restrict__ corr_out, float* __restrict__ rec_out) {
    __shared__ __attribute__((aligned(16))) float s_all[4][256 + 232 + 240];
    const CorrArgs& A = *args_p;
    const RigArgs& R = *rig_p;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int t = xcd_block(blockIdx.x, gridDim.x) * 4 + wave;
    if (t >= num_tasks) return;   // wave-uniform: no block barriers below
    float* S = s_all[wave];
    float* s_corr = S + 256;
    float* s_combo = s_corr + 232;
    const float qnan = __int_as_float(0x7fc00000);
    float* rec = rec_out + (size_t)t * RIG_REC;
    if (tasks[(size_t)t * task_words] == 0) {
        if (FROM_MIX)
            for (int o = lane; o < CLT_CORR_LEN; o += 64) corr_out[(size_t)t * CLT_CORR_LEN + o] = qnan;
        if (lane < RIG_REC) rec[lane] = (lane == 1) ? 0.f : qnan;
        return;
    }
    if constexpr (FROM_MIX) {
        float D1[3][4], D2[3][4];
#pragma unroll
        for (int col = 0; col < 3; col++)
#pragma unroll
            for (int n = 0; n < 4; n++) {
                D1[col][n] = mix_main[((size_t)t * 3 + col) * 256 + n * 64 + lane];
                D2[col][n] = mix_aux[((size_t)t * 3 + col) * 256 + n * 64 + lane];
            }
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int col = 0; col < 3; col++) {   // correlateSingleColorFD (C2:197-231), colour mix (C2:275-310)
            const float* d1 = D1[col];
            const float* d2 = D2[col];
            const float n1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2] + d1[3] * d1[3];
            const float n2 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2] + d2[3] * d2[3];
            const float a2 = fast_sqrt(n1 * n2);
            float fz2 = 0.f;
            if (FZ) fz2 = wave_sum(a2) * (1.f / 64.f) * A.fat_zero * A.fat_zero;
            const float scale = fast_rcp(a2 + fz2) * A.col_w[col];
            const float t0 = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2] + d1[3] * d2[3];
            const float t1 = -d1[1] * d2[0] + d1[0] * d2[1] - d1[3] * d2[2] + d1[2] * d2[3];
            const float t2 = -d1[2] * d2[0] - d1[3] * d2[1] + d1[0] * d2[2] + d1[1] * d2[3];
            const float t3 = d1[3] * d2[0] - d1[2] * d2[1] - d1[1] * d2[2] + d1[0] * d2[3];
            acc[0] = fmaf(t0, scale, acc[0]);
            acc[1] = fmaf(t1, scale, acc[1]);
            acc[2] = fmaf(t2, scale, acc[2]);
            acc[3] = fmaf(t3, scale, acc[3]);
        }
        float umx[4];
        corr_mx_operand(lane, umx);
        corr_fd_to_tile_mx(acc, lpf[lane], S, umx, lane, corr_out + (size_t)t * CLT_CORR_LEN, s_corr);
    } else {
        for (int o = lane; o < CLT_CORR_LEN; o += 64) s_corr[o] = src[(size_t)t * CLT_CORR_LEN + o];
    }
    wave_sync();
    // the single correlation as a horizontal pair's strip (scaleRotateInterpoateCorrelations with
    // pair mask 1) and combineInterpolatedCorrelations of that one strip (C2:730-790)
    const int hw = A.strip_height, slen = hw * 15;
    for (int e = lane; e < slen; e += 64) {
        const int row = e / 15, scol = e - row * 15 - 7;
        const StripAt<0, 1> sh(row, scol);
        const float st = sh(s_corr);
        float v;
        if (A.corr_offset >= 0.f) {
            float dd = st + A.corr_offset;
            if (dd < 0.f) dd = 0.f;
            if (A.twice_diagonal) {
                v = dd * dd;
                if (v > 0.f) v = fast_sqrt(v) - A.corr_offset;
            } else {
                v = dd;
                if (v > 0.f) v -= A.corr_offset;
            }
        } else {
            v = st;
        }
        s_combo[e] = sh.nan ? qnan : v;
    }
    wave_sync();
    // integer maximum on the axis (getMaxXYInt, C2:802-843)
    const int imx = __shfl(argmax16(lane < 15 ? s_combo[lane & 15] : -INFINITY, lane & 15), 0);
    const float vmax = s_combo[imx];
    const int ix = imx - 7;
    float o_disp = qnan, o_str = 0.f, o_x0 = qnan, cmdx = qnan, cmdy = qnan;
    if (vmax >= R.min_corr) {
        o_disp = (float)(-ix);
        o_str = vmax;
        // centre of mass along the strip rows (getMaxXCm / getMaxXCmNotch, C2:925-1032): row dy
        // reads adx = dy & 1, dy & 1 + 2, ... < 9 on both sides
        const int ixc = R.cm_at_zero ? 0 : ix, x0 = 7 + ixc;
        float s0 = 0.f, sx = 0.f, sx2 = 0.f;
        const int rows = R.cm_rows;
        for (int k = lane; k < 16 * 9; k += 64) {
            const int dy = k / 9, j = k - dy * 9;   // j: sample of the row (0..8 even rows, 0..7 odd rows)
            if (dy >= rows) continue;
            const int odd = dy & 1;
            int adx, dir;
            if (!odd) {
                adx = (j + 1) & ~1;
                dir = (j == 0 || (j & 1) == 0) ? 1 : -1;
            } else {
                if (j >= 8) continue;
                adx = 2 * (j >> 1) + 1;
                dir = (j & 1) ? 1 : -1;
            }
            const int idx = (adx * dir) >> 1;
            const int xq = 2 * idx + odd, x1 = x0 + idx;
            if (x1 < 0 || x1 >= 15) continue;
            float d = s_combo[15 * dy + x1];
            if (isnan(d) || !(d > 0.f)) continue;
            d *= R.cm_wy[dy] * R.wndx[adx];
            s0 += d;
            sx += d * xq;
            sx2 += d * xq * xq;
        }
        s0 = wave_sum(s0);
        sx = wave_sum(sx);
        sx2 = wave_sum(sx2);
        if (s0 != 0.f) {
            const float cs0 = ixc + sx / s0 / 2.f;
            const float cs2 = fast_sqrt(s0 * sx2 - sx * sx) / s0 / 2.f;
            const float eff = cs2 * R.cm_max_norm;
            o_disp = -cs0;
            o_str = s0 / (eff * eff);
            o_x0 = cs0;
        }
        // single2dCM around (0, 0) (C2:1538-1586): positive values within the radius
        float c0 = 0.f, cx = 0.f, cy = 0.f;
        const int ir = R.iradius, side = 2 * ir + 1;
        for (int k = lane; k < side * side; k += 64) {
            const int yy = k / side - ir, xv = k - (k / side) * side - ir;
            if (yy < -7 || yy > 7 || xv < -7 || xv > 7) continue;
            if (!R.square && yy * yy + xv * xv > R.ir2) continue;
            const float d = s_corr[(yy + 7) * 15 + xv + 7];
            if (d > 0.f) {
                c0 += d;
                cx += d * xv;
                cy += d * yy;
            }
        }
        c0 = wave_sum(c0);
        cx = wave_sum(cx);
        cy = wave_sum(cy);
        cmdx = -(cx / c0);
        cmdy = -(cy / c0);
    }
    if (lane < RIG_REC) {
        float v = qnan;
        switch (lane) {
        case 0: v = o_disp; break;
        case 1: v = o_str; break;
        case 2: v = o_x0; break;
        case 4: v = cmdx; break;
        case 5: v = cmdy; break;
        default: break;
        }
        rec[lane] = v;
    }
}

// Low-texture pass (ID:9404-9454): the inter correlations of the (2 lt_rad + 1) x (2 lt_rad_x + 1)
// neighbouring task tiles, weighted by cos(pi |dy| / (2 lt_rad + 1)) cos(pi |dx| / (2 lt_rad + 1)),
// normalised by the sum of the weights.  One wavefront per task.
__global__ __launch_bounds__(256) void k_inter_lt_avg(const float* __restrict__ corr, const int32_t* __restrict__ tasks,
                                                      int task_words, int num_tasks, const int32_t* __restrict__ tile_map,
                                                      int lt_rad, int lt_rad_x, GeomDev g, float* __restrict__ out) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int t = blockIdx.x * 4 + wave;
    if (t >= num_tasks) return;
    const int32_t* rec = tasks + (size_t)t * task_words;
    float* o = out + (size_t)t * CLT_CORR_LEN;
    const float qnan = __int_as_float(0x7fc00000);
    if (rec[0] == 0) {
        for (int e = lane; e < CLT_CORR_LEN; e += 64) o[e] = qnan;
        return;
    }
    const int tx = rec[1] & 0xffff, ty = (rec[1] >> 16) & 0xffff;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    float sw = 0.f;
    const float k = (float)M_PI / (2 * lt_rad + 1);
    for (int dy = -lt_rad; dy <= lt_rad; dy++) {
        const int y = ty + dy;
        if (y < 0 || y >= g.tiles_y) continue;
        for (int dx = -lt_rad_x; dx <= lt_rad_x; dx++) {
            const int x = tx + dx;
            if (x < 0 || x >= g.tiles_x) continue;
            const int n = tile_map[y * g.tiles_x + x];
            if (n < 0 || tasks[(size_t)n * task_words] == 0) continue;   // tiles without a correlation
            const float nw = cosf(k * abs(dy)) * cosf(k * abs(dx));
            const float* c = corr + (size_t)n * CLT_CORR_LEN;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int e = lane + 64 * i;
                if (e < CLT_CORR_LEN) acc[i] = fmaf(nw, c[e], acc[i]);
            }
            sw += nw;
        }
    }
    const float s = 1.f / sw;   // the tile itself is always in the map
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int e = lane + 64 * i;
        if (e < CLT_CORR_LEN) o[e] = acc[i] * s;
    }
}

// The BIDISPARITY record (ID:89-114, SNGL_TO_BI ID:190-194, ID:9186-9248) of every task from the
// two quads' records and one-direction fits and the inter result; skipped tasks: zeros and a NaN
// target (the reference's zero-initialised arrays, ID:8964-8969).  extra_main / extra_aux: each
// quad's disparity, the extra disparity of its textures (ID:9194-9197).
__global__ __launch_bounds__(256) void k_rig_bimap(const int32_t* __restrict__ tasks, int task_words, int num_tasks,
                                                   const float* __restrict__ disp_main, const float* __restrict__ dir_main,
                                                   const float* __restrict__ disp_aux, const float* __restrict__ dir_aux,
                                                   const float* __restrict__ inter, const float* __restrict__ target,
                                                   float* __restrict__ bimap, float* __restrict__ extra_main,
                                                   float* __restrict__ extra_aux) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_tasks) return;
    float* b = bimap + (size_t)t * CLT_BI_LAYERS;
    const float qnan = __int_as_float(0x7fc00000);
    float v[CLT_BI_LAYERS];
#pragma unroll
    for (int i = 0; i < CLT_BI_LAYERS; i++) v[i] = 0.f;
    if (tasks[(size_t)t * task_words] == 0) {
        v[CLT_BI_TARGET] = qnan;
    } else {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const float* r = (q ? disp_aux : disp_main) + (size_t)t * CLT_DISP_LAYERS;
            const float* d = (q ? dir_aux : dir_main) + (size_t)t * 8;
            const int bd = q ? CLT_BI_ADISP_FULL : CLT_BI_DISP_FULL, bs = q ? CLT_BI_ASTR_FULL : CLT_BI_STR_FULL;
            const float cm = r[CLT_DISP_CM];
            v[bd] = isnan(cm) ? r[CLT_DISP_INT] : cm;   // -ixy when the centre of mass has no weight
            v[bs] = r[CLT_DISP_STRENGTH];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                v[bd + 1 + k] = d[k];
                v[bs + 1 + k] = d[4 + k];
            }
        }
        const float* x = inter + (size_t)t * 4;
        v[CLT_BI_DISP_CROSS] = x[0];
        v[CLT_BI_STR_CROSS] = x[1];
        v[CLT_BI_DISP_CROSS_DX] = x[2];
        v[CLT_BI_DISP_CROSS_DY] = x[3];
        v[CLT_BI_STR_ALL] = (float)pow((double)x[1] * v[CLT_BI_STR_FULL] * v[CLT_BI_ASTR_FULL], 1.0 / 3);
        v[CLT_BI_TARGET] = target ? target[t] : qnan;
    }
#pragma unroll
    for (int i = 0; i < CLT_BI_LAYERS; i++) b[i] = v[i];
    extra_main[t] = v[CLT_BI_DISP_FULL];
    extra_aux[t] = v[CLT_BI_ADISP_FULL];
}

}  // namespace clt
