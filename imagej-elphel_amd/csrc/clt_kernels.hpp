// clt_kernels.hpp -- the four hot kernels of the tile processor and the texture combiner.
#pragma once
#include "clt_device.hpp"

namespace clt {

// ====================================================================================
// convert_direct: Bayer/mono mosaic -> per tile/camera/colour 16x16 lapped DCT-IV/DST-IV
// with aberration-kernel multiply and fractional shift (ImageDtt.java:4867-5584 semantics;
// replaces TileProcessor.cuh convert_correct_tiles).
// Unit u = (task, camera, colour); 16 lanes per unit, 16 units per 256-thread block.
// ====================================================================================
#pragma clang fp contract(off)
// Integer window corner and residual shift of one unit, in double precision with the operation
// order of ImageDtt.java:4904-4943 so the integer corner is bit-identical to the CPU reference.
struct TilePos {
    int left, top;
    float rx, ry;
};
__device__ __forceinline__ TilePos tile_position(double cx, double cy, const float* ce, int kstep, int ktx, int kty) {
    double px = cx - 8.0, py = cy - 8.0;
    if (ce) {
        double kdx = cx - ((double)(ktx - 1) + 0.5) * kstep;
        double kdy = cy - ((double)(kty - 1) + 0.5) * kstep;
        px = cx - 8.0 - ((double)ce[0] + (double)ce[4] * kdx + (double)ce[5] * kdy);
        py = cy - 8.0 - ((double)ce[1] + (double)ce[6] * kdx + (double)ce[7] * kdy);
    }
    TilePos p;
    p.left = (int)-floor(-px + 0.5);
    p.top = (int)-floor(-py + 0.5);
    p.rx = (float)(-(px - p.left));
    p.ry = (float)(-(py - p.top));
    return p;
}
#pragma clang fp contract(fast)

__device__ __forceinline__ int clamp_parity(int p, int n) {   // ImageDtt.java:4958-4965
    if (p < 0) return p & 1;
    if (p >= n) return n - 2 + (p & 1);
    return p;
}


// LDS of one 16-lane unit: 16 rows of CONV_ROW floats (16-byte row accesses of 8 consecutive lanes
// cover all banks once), units CONV_UNIT floats apart (neighbouring units on the other half of the
// banks).
constexpr int CONV_ROW = 20;
constexpr int CONV_UNIT = 336;

// Block = `colors` wavefronts, wavefront w converts colour w of 4 (task, camera) units, 16 lanes
// per unit.  Every row of a Bayer colour plane holds samples of one column parity only, so its
// sine fold transform is the reversed cosine one: DST-IV(fold_S x)[k] = s * DCT-IV(fold_C x)[7-k],
// s = -1 when the samples sit at odd window columns (same along columns for red and blue).
// The block's work; returns false when the wavefront has no active unit (it stored nothing), t_out
// = the task of the wavefront's first unit.  On return true the units' CLT is also in s_buf
// ([unit][mode * 64 + kh * 8 + kv], unit = colour * 4 + unit of the wavefront).
__device__ __forceinline__ bool convert_tiles(const float* __restrict__ images, const float* __restrict__ kernels,
                                              const float* __restrict__ koffsets, const int32_t* __restrict__ tasks,
                                              int task_words, int num_tasks, float* __restrict__ clt,
                                              const float* __restrict__ lpf, int lpf_mask, double inv_kstep,
                                              const int32_t* __restrict__ task_frame, long long img_fstride,
                                              long long clt_fstride, const GeomDev& g, float (*s_buf)[CONV_UNIT],
                                              int& t_out) {
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int color = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform
    const int unit_in_block = threadIdx.x >> 4, l = threadIdx.x & 15;
    const int total = num_tasks * g.cams;
    int u = blk * 4 + ((threadIdx.x >> 4) & 3);             // (task, camera) unit
    bool active = u < total;
    if (!active) u = 0;
    const int t = band_order(u / g.cams, num_tasks, g.tiles_x), cam = u % g.cams;
    t_out = t;
    const int32_t* rec = tasks + (size_t)t * task_words;
    active = active && (rec[0] != 0);
    if (!__any(active)) return false;   // whole wave idle (no block-level barriers here)
    const long long frame = (task_frame && active) ? task_frame[t] : 0;   // batched frames
    const bool bayer = g.colors == 3;
    const int txy = rec[1];
    const int tx = txy & 0xffff, ty = (txy >> 16) & 0xffff;
    const double cx = active ? (double)__int_as_float(rec[2 + 2 * cam]) : 8.0;
    const double cy = active ? (double)__int_as_float(rec[3 + 2 * cam]) : 8.0;

    // ---- kernel tile (ImageDtt.java:4904-4911) and integer/residual position
    const float* ce = nullptr;
    const float* kt = nullptr;
    int ktx = 0, kty = 0;
    if (kernels) {
        ktx = (int)floor(cx * inv_kstep + 0.5) + 1;        // exact: kstep is a power of two
        kty = (int)floor(cy * inv_kstep + 0.5) + 1;
        kty = kty < 0 ? 0 : (kty >= g.kv ? g.kv - 1 : kty);
        ktx = ktx < 0 ? 0 : (ktx >= g.kh ? g.kh - 1 : ktx);
        const size_t kidx = (((size_t)cam * g.kv + kty) * g.kh + ktx) * g.colors + color;
        ce = koffsets + kidx * 8;
        kt = kernels + kidx * 256;
    }
    // aberration kernel, loaded early so the load overlaps the window loads and the transforms:
    // the four quadrants of row kh = l & 7 at kv = 4 hs .. 4 hs + 3, hs = l >> 3 (the kv half this
    // lane takes in the frequency-domain stages below)
    float kk[4][4];
    if (kt) {
        const float* kq = kt + (l & 7) * 8 + 4 * (l >> 3);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float4 v = *reinterpret_cast<const float4*>(kq + 64 * q);
            kk[q][0] = v.x; kk[q][1] = v.y; kk[q][2] = v.z; kk[q][3] = v.w;
        }
    }
    const TilePos pos = tile_position(cx, cy, ce, g.kstep, ktx, kty);
    const int left = pos.left, top = pos.top;
    const float rx = pos.rx, ry = pos.ry;

    // ---- 16x16 window: coalesced row loads (lane l = window column l, one row per load), staged
    // through LDS so that lane r then holds window row r.  Pixels of other colours are 0; outside
    // the frame the parity-preserving clamp of ImageDtt.java:4958-4969 applies.
    float* sb = s_buf[unit_in_block];   // [16][CONV_ROW] floats, reused by the stage H output below
    {
        const float* img = images + frame * img_fstride + (size_t)cam * g.height * g.width;
        const int pc = left + l;
        const int cp = pc & 1;
        // keep pixel (row parity rp, column parity cp) of this colour
        bool keep_e, keep_o;   // rows top + even / top + odd
        {
            const int rp_e = top & 1, rp_o = rp_e ^ 1;
            auto colour_at = [&](int rp) {
                const bool rr = rp == g.red_row, rc = cp == g.red_col;
                return rr ? (rc ? 0 : 2) : (rc ? 2 : 1);   // 0 red, 1 blue, 2 green
            };
            keep_e = active && (!bayer || colour_at(rp_e) == color);
            keep_o = active && (!bayer || colour_at(rp_o) == color);
        }
        const bool inside = (left >= 0) && (left + 16 <= g.width) && (top >= 0) && (top + 16 <= g.height);
        if (__all(inside || !active)) {
            // (an inactive lane reads pixel (0, 0))
            const float* src = img + (active ? (size_t)top * g.width + pc : 0);
            const size_t stride = active ? (size_t)g.width : 0;
            if (bayer && color != 2) {   // red / blue: only the 8 rows of this colour's parity are read
                const int rsel = (((color == 0) ? g.red_row : (g.red_row ^ 1)) ^ top) & 1;
                const bool keep = rsel ? keep_o : keep_e;
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int rr = 2 * i + rsel;
                    const float v = src[rr * stride];
                    sb[rr * CONV_ROW + l] = keep ? v : 0.f;
                    sb[(rr ^ 1) * CONV_ROW + l] = 0.f;
                }
            } else {
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const float v = src[r * stride];
                    sb[r * CONV_ROW + l] = ((r & 1) ? keep_o : keep_e) ? v : 0.f;
                }
            }
        } else {
            const int pcc = clamp_parity(pc, g.width);
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int pr = clamp_parity(top + r, g.height);
                const bool keep = (r & 1) ? keep_o : keep_e;
                sb[r * CONV_ROW + l] = keep ? img[(size_t)pr * g.width + pcc] : 0.f;
            }
        }
    }
    wave_sync();
    const int r = l;
    const int prow = top + r;
    const bool row_red = ((prow & 1) == g.red_row);
    // column parity (absolute) holding this colour on this row; -1: none (empty R/B row)
    int cpar;
    if (!bayer) cpar = 2;                                              // mono: every column
    else if (color == 0) cpar = row_red ? g.red_col : -1;
    else if (color == 1) cpar = row_red ? -1 : (g.red_col ^ 1);
    else cpar = row_red ? (g.red_col ^ 1) : g.red_col;
    float x[16];
#pragma unroll
    for (int j = 0; j < 16; j += 4) {
        const float4 v = *reinterpret_cast<const float4*>(&sb[r * CONV_ROW + j]);
        x[j] = v.x; x[j + 1] = v.y; x[j + 2] = v.z; x[j + 3] = v.w;
    }
    wave_sync();
    float ahs, ahc, avs, avc;
    sincos_pi(rx * 0.0625f, &ahs, &ahc);
    sincos_pi(ry * 0.0625f, &avs, &avc);
    {   // ---- stage H: horizontal fold + DCT-IV (C) / DST-IV (S)
        float yc[8], ys[8], hc[8], hsn[8];
        fold16(x, ahc, ahs, yc, ys);
        dct_iv8(yc, hc);
        if (bayer) {
            const float sg = (((cpar ^ left) & 1) ? -1.f : 1.f);        // odd window columns: -1
#pragma unroll
            for (int k = 0; k < 8; k++) hsn[k] = sg * hc[7 - k];
        } else {
            dst_iv8(ys, hsn);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            sb[k * CONV_ROW + r] = hc[k];
            sb[(8 + k) * CONV_ROW + r] = hsn[k];
        }
    }
    wave_sync();

    // ---- stage V: lane = (hs, kh); vertical fold + DCT-IV (mode hs) / DST-IV (mode hs|2)
    const int hs = l >> 3, kh = l & 7;
    float col[16];
#pragma unroll
    for (int i = 0; i < 16; i += 4) {
        float4 v = *reinterpret_cast<const float4*>(&sb[(hs * 8 + kh) * CONV_ROW + i]);
        col[i] = v.x; col[i + 1] = v.y; col[i + 2] = v.z; col[i + 3] = v.w;
    }
    float a[8], b[8];
    {
        float zc[8], zs[8];
        fold16(col, avc, avs, zc, zs);
        dct_iv8(zc, a);    // mode hs
        if (bayer && color != 2) {   // red / blue: rows of one parity only (wave-uniform branch)
            const int rpar = (color == 0) ? g.red_row : (g.red_row ^ 1);
            const float sg = (((rpar ^ top) & 1) ? -1.f : 1.f);
#pragma unroll
            for (int k = 0; k < 8; k++) b[k] = sg * a[7 - k];
        } else {
            dst_iv8(zs, b);   // mode hs | 2
        }
    }
    const float sigma = hs ? 1.f : -1.f;

    // ---- frequency-domain stages: aberration kernel multiply (ImageDtt.java:5429-5488), fractional
    // shift (ImageDtt.java:5490-5584) and LPF.  Each mixes the modes of the partner lanes (hs = 0, 1
    // of row kh) at equal kv, so the two lanes split the work by kv: lane hs takes kv = 4 hs + j,
    // j = 0..3, of both lanes' modes.  One exchange gives each lane the partner's a, b of its half;
    // oa, ob are the lane's own modes (hs, hs | 2), pa, pb the partner's (hs ^ 1, (hs ^ 1) | 2).
    float oa[4], ob[4], pa[4], pb[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        pa[j] = lane_xor8(hs ? a[j] : a[j + 4]);   // the partner sends what this lane takes
        pb[j] = lane_xor8(hs ? b[j] : b[j + 4]);
        oa[j] = hs ? a[j + 4] : a[j];
        ob[j] = hs ? b[j + 4] : b[j];
    }
    // with sign sigma of the lane's own modes, the partner's sign is -sigma
    if (kt) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float k0 = kk[0][j], k1 = kk[1][j], k2 = kk[2][j], k3 = kk[3][j];
            const float spa = sigma * pa[j], spb = sigma * pb[j], soa = sigma * oa[j], sob = sigma * ob[j];
            const float na = oa[j] * k0 + spa * k1 - ob[j] * k2 - spb * k3;
            const float nb = ob[j] * k0 + spb * k1 + oa[j] * k2 + spa * k3;
            const float npa = pa[j] * k0 - soa * k1 - pb[j] * k2 + sob * k3;
            const float npb = pb[j] * k0 - sob * k1 + pa[j] * k2 - soa * k3;
            oa[j] = na; ob[j] = nb; pa[j] = npa; pb[j] = npb;
        }
    }
    {   // horizontal shift by rx (kh), vertical by ry (kv = 4 hs + j)
        float sh, ch, cv, sv, c1, s1;
        sincos_pi((kh + 0.5f) * rx * 0.125f, &sh, &ch);
        sincos_pi((4 * hs + 0.5f) * ry * 0.125f, &sv, &cv);   // then rotations by pi ry / 8
        sincos_pi(ry * 0.125f, &s1, &c1);
        const float ssh = sigma * sh;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float ta = oa[j] * ch + pa[j] * ssh, tb = ob[j] * ch + pb[j] * ssh;
            const float tpa = pa[j] * ch - oa[j] * ssh, tpb = pb[j] * ch - ob[j] * ssh;
            oa[j] = ta * cv - tb * sv;
            ob[j] = tb * cv + ta * sv;
            pa[j] = tpa * cv - tpb * sv;
            pb[j] = tpb * cv + tpa * sv;
            const float cn = fmaf(cv, c1, -sv * s1), sn = fmaf(sv, c1, cv * s1);
            cv = cn;
            sv = sn;
        }
    }
    if (lpf && ((lpf_mask >> color) & 1)) {
        const float4 f = *reinterpret_cast<const float4*>(lpf + color * 64 + kh * 8 + 4 * hs);
        const float fv[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            oa[j] *= fv[j];
            ob[j] *= fv[j];
            pa[j] *= fv[j];
            pb[j] *= fv[j];
        }
    }
    // ---- store through the unit's LDS: each lane's four kv-halves go to their places in the
    // unit's 256 coefficients (the two lanes of a pair write disjoint banks), then lane l stores
    // coefficients 4l .. 4l + 3 of each mode, so every store instruction writes whole 256-byte runs
    wave_sync();   // stage V's reads of sb are done
    {
        const int o = kh * 8 + 4 * hs;
        *reinterpret_cast<float4*>(&sb[hs * 64 + o]) = make_float4(oa[0], oa[1], oa[2], oa[3]);
        *reinterpret_cast<float4*>(&sb[(hs | 2) * 64 + o]) = make_float4(ob[0], ob[1], ob[2], ob[3]);
        *reinterpret_cast<float4*>(&sb[(hs ^ 1) * 64 + o]) = make_float4(pa[0], pa[1], pa[2], pa[3]);
        *reinterpret_cast<float4*>(&sb[((hs ^ 1) | 2) * 64 + o]) = make_float4(pb[0], pb[1], pb[2], pb[3]);
    }
    wave_sync();
    if (active) {
        const size_t tile = (size_t)ty * g.tiles_x + tx;
        float* o = clt + frame * clt_fstride + (((size_t)cam * g.tiles_x * g.tiles_y + tile) * g.colors + color) * 256 + 4 * l;
#pragma unroll
        for (int m = 0; m < 4; m++)
            *reinterpret_cast<float4*>(o + m * 64) = *reinterpret_cast<const float4*>(&sb[m * 64 + 4 * l]);
    }
    return true;
}
__global__ __launch_bounds__(192) void k_convert(const float* __restrict__ images, const float* __restrict__ kernels,
                                                 const float* __restrict__ koffsets, const int32_t* __restrict__ tasks,
                                                 int task_words, int num_tasks, float* __restrict__ clt,
                                                 const float* __restrict__ lpf, int lpf_mask, double inv_kstep,
                                                 const int32_t* __restrict__ task_frame, long long img_fstride,
                                                 long long clt_fstride, GeomDev g) {
    __shared__ __attribute__((aligned(16))) float s_buf[12][CONV_UNIT];
    int t;
    (void)convert_tiles(images, kernels, koffsets, tasks, task_words, num_tasks, clt, lpf, lpf_mask, inv_kstep,
                        task_frame, img_fstride, clt_fstride, g, s_buf, t);
}

// ====================================================================================
// Batched frames: the task lists of the frames (staged frame-major, max_tasks records apart) are
// packed into one list with the frame of each task, and the per-frame tile maps are rebuilt.
// ====================================================================================
__global__ __launch_bounds__(256) void k_fill_i32(int32_t* __restrict__ p, long long n, int32_t v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}

__global__ __launch_bounds__(256) void k_batch_assemble(const int32_t* __restrict__ staged, const int32_t* __restrict__ counts,
                                                        const int32_t* __restrict__ starts, int num_frames, int max_tasks,
                                                        int task_words, int32_t* __restrict__ tasks,
                                                        int32_t* __restrict__ task_frame, int32_t* __restrict__ tile_map,
                                                        GeomDev g) {
    const long long slot = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (slot >= (long long)num_frames * max_tasks) return;
    const int f = (int)(slot / max_tasks), i = (int)(slot % max_tasks);
    if (i >= counts[f]) return;
    const int dst = starts[f] + i;
    const int32_t* src = staged + slot * task_words;
    int32_t* d = tasks + (size_t)dst * task_words;
    for (int w = 0; w < task_words; w++) d[w] = src[w];
    task_frame[dst] = f;
    if (src[0] != 0) {
        const int tx = src[1] & 0xffff, ty = (src[1] >> 16) & 0xffff;
        tile_map[(size_t)f * g.tiles_x * g.tiles_y + (size_t)ty * g.tiles_x + tx] = dst;
    }
}

// ====================================================================================
// Port geometry (GeometryCorrection.getPortsCoordinatesAndDerivatives, GC:2159-2314) in double
// precision, one thread per (task, sensor); replaces the host loop of getTileSubcamOffsets
// (GTP:686-778).  Contraction off: the same rounding as the double-precision reference.
// ====================================================================================
struct PortGeomDev {
    double rots[4][9];          // per sensor rotation/zoom matrix (GC:1043-1076)
    double drots[4][4][9];      // its derivatives by azimuth, tilt, roll, zoom (GC:1083-1147)
    double rxy[4][2], pxy0[4][2];
    double dist[7];             // C, B, A, A5, A6, A7, A8
    double fl_pix, ri_scale, half_w, half_h, rd_scale, step_r;
    int tab_len;
};

#pragma clang fp contract(off)
// Port centre (x, y) of sensor i for tile centre (px, py) at disparity d; with rx, ry non-null
// also its derivatives: 14 values each, by the 13 extrinsic parameters and by the disparity.
__device__ __forceinline__ void port_centre(const PortGeomDev& P, const double* __restrict__ rtab, double px, double py,
                                            double d, int i, double* x, double* y, double* rx, double* ry) {
    // undo the common distortion of the tile centre (GC:2181-2187, getRByRDist GC:2802-2830)
    const double pXcd = px - P.half_w, pYcd = py - P.half_h;
    const double rrel = sqrt(pXcd * pXcd + pYcd * pYcd) * P.rd_scale;
    double rND2R = __longlong_as_double(0x7ff8000000000000LL);
    if (rrel >= 0.0) {
        const int index = (int)floor(rrel / P.step_r);
        if (index < P.tab_len - 1) rND2R = rtab[index] + (rtab[index + 1] - rtab[index]) * (rrel / P.step_r - index);
    }
    const double vi[3] = {pXcd * rND2R - d * P.rxy[i][0], pYcd * rND2R - d * P.rxy[i][1], P.fl_pix};
    const double* R = P.rots[i];
    const double rvi[3] = {R[0] * vi[0] + R[1] * vi[1] + R[2] * vi[2], R[3] * vi[0] + R[4] * vi[1] + R[5] * vi[2],
                           R[6] * vi[0] + R[7] * vi[1] + R[8] * vi[2]};
    const double norm_z = P.fl_pix / rvi[2];
    const double pXci = rvi[0] * norm_z, pYci = rvi[1] * norm_z;
    const double rNDi = sqrt(pXci * pXci + pYci * pYci);
    const double ri = rNDi * P.ri_scale;
    double k = 1.0, rri = 1.0, dk_dri = 0.0;
#pragma unroll
    for (int j = 0; j < 7; j++) {   // GC:2217-2222 and GC:2264-2268
        dk_dri += P.dist[j] * (j + 1) * rri;
        rri *= ri;
        k += P.dist[j] * (rri - 1.0);
    }
    *x = pXci * k + P.pxy0[i][0];
    *y = pYci * k + P.pxy0[i][1];
    if (!rx) return;
    double dx[5], dy[5];   // by azimuth, tilt, roll, zoom (GC:2235-2283), disparity
#pragma unroll
    for (int a = 0; a < 5; a++) {
        double dv[3];
        if (a < 4) {
            const double* D = P.drots[i][a];
#pragma unroll
            for (int r = 0; r < 3; r++) dv[r] = D[r * 3] * vi[0] + D[r * 3 + 1] * vi[1] + D[r * 3 + 2] * vi[2];
        } else {
#pragma unroll
            for (int r = 0; r < 3; r++) dv[r] = -(R[r * 3] * P.rxy[i][0] + R[r * 3 + 1] * P.rxy[i][1]);
        }
        const double dxc = dv[0] * norm_z - pXci * dv[2] / rvi[2];
        const double dyc = dv[1] * norm_z - pYci * dv[2] / rvi[2];
        const double dri = (a == 2) ? 0.0 : P.ri_scale / rNDi * (pXci * dxc + pYci * dyc);   // roll keeps the radius
        const double dkk = dk_dri * dri;
        dx[a] = dxc * k + pXci * dkk;
        dy[a] = dyc * k + pYci * dkk;
    }
    for (int c = 0; c < 14; c++) rx[c] = ry[c] = 0.0;
    if (i < 3) {   // sensor 3 tilt/azimuth/zoom are minus the sum of the others (GC:2285-2302)
        rx[0 + i] = dx[1]; ry[0 + i] = dy[1];
        rx[3 + i] = dx[0]; ry[3 + i] = dy[0];
        rx[10 + i] = dx[3]; ry[10 + i] = dy[3];
    } else {
        for (int j = 0; j < 3; j++) {
            rx[0 + j] = -dx[1]; ry[0 + j] = -dy[1];
            rx[3 + j] = -dx[0]; ry[3 + j] = -dy[0];
            rx[10 + j] = -dx[3]; ry[10 + j] = -dy[3];
        }
    }
    rx[6 + i] = dx[2]; ry[6 + i] = dy[2];
    rx[13] = dx[4]; ry[13] = dy[4];
}

// rig: pg[0] the main quad (sensors 0..3), pg[1] the auxiliary quad (sensors 4..7, clt_bi_quad
// ID:8970-8993) at disparity * dscale (R_aux / R_main); derivatives only for a single quad.
__global__ __launch_bounds__(256) void k_port_centres(const PortGeomDev* __restrict__ pg, const double* __restrict__ rtab,
                                                      const int32_t* __restrict__ words, const int32_t* __restrict__ txy,
                                                      const float* __restrict__ disparity, int num_tasks, int task_words,
                                                      int32_t* __restrict__ tasks, double* __restrict__ deriv, int rig,
                                                      double dscale) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int shift = rig ? 3 : 2;
    const int t = gid >> shift, i = gid & ((1 << shift) - 1);
    const int q = i >> 2, s = i & 3;
    if (t >= num_tasks) return;
    int32_t* rec = tasks + (size_t)t * task_words;
    const int word = words[t], xy16 = txy[t];
    if (i == 0) {
        rec[0] = word;
        rec[1] = xy16;
    }
    double* dr = (deriv && !rig) ? deriv + (size_t)t * 8 * 14 + (size_t)(2 * i) * 14 : nullptr;
    if (word == 0) {   // GTP:720-722: nothing to do for this tile
        rec[2 + 2 * i] = 0;
        rec[3 + 2 * i] = 0;
        if (dr)
            for (int k = 0; k < 28; k++) dr[k] = 0.0;
        return;
    }
    const double px = (double)((xy16 & 0xffff) * 8 + 4), py = (double)(((xy16 >> 16) & 0xffff) * 8 + 4);   // GTP:724-725
    double x, y;
    port_centre(pg[q], rtab, px, py, (double)disparity[t] * (q ? dscale : 1.0), s, &x, &y, dr, dr ? dr + 14 : nullptr);
    rec[2 + 2 * i] = __float_as_int((float)x);
    rec[3 + 2 * i] = __float_as_int((float)y);
}

#pragma clang fp contract(fast)

// ====================================================================================
// imclt_rbg: inverse CLT + overlap-add, gather form (no read-modify-write, one launch).
// Block = strip of IMCLT_S output 8x8 blocks wide, IMCLT_R block rows tall; it sweeps down
// the tile rows keeping the bottom halves of the previous tile row in LDS.  Output image of a
// camera/colour is (H+8) x (W+8), uncropped (GPUTileProcessor.java:599-626).
// ====================================================================================
constexpr int IMCLT_S = 15;   // output blocks per strip (16 tiles incl. the left halo)

// grid: (strips, segments of seg_rows block rows, frames * cams * colours).  A segment without any
// task tile only writes its zeros.
__global__ __launch_bounds__(256) void k_imclt(const float* __restrict__ clt, const int32_t* __restrict__ tile_map,
                                               const float* __restrict__ lpf, float* __restrict__ rbg, int seg_rows,
                                               long long clt_fstride, GeomDev g) {
    // per unit: IMCLT scratch, then (after the unit's last scratch read) its 16x16 tile, row stride 17
    __shared__ __attribute__((aligned(16))) float s_scr[16][IMCLT_SCR_UNIT];
    __shared__ float s_carry[8][IMCLT_S * 8];
    static_assert(16 * 17 <= IMCLT_SCR_UNIT, "tile must fit the unit scratch");
    auto s_tile = [&](int u, int r, int c) -> float& { return s_scr[u][r * 17 + c]; };
    const int planes = g.cams * g.colors;
    const int frame = blockIdx.z / planes, plane = blockIdx.z % planes;   // plane = cam * colors + color
    const int cam = plane / g.colors, color = plane % g.colors;
    const int bx0 = blockIdx.x * IMCLT_S, by0 = blockIdx.y * seg_rows;
    const int out_w = g.width + 8, out_h = g.height + 8;
    const int nby = min(seg_rows, g.tiles_y + 1 - by0);
    const int unit = threadIdx.x >> 4, l = threadIdx.x & 15;
    const int hs = l >> 3, kh = l & 7;
    const int tx = bx0 - 1 + unit;
    const size_t ntiles = (size_t)g.tiles_x * g.tiles_y;
    tile_map += (size_t)frame * ntiles;
    clt += frame * clt_fstride;
    float* out = rbg + ((size_t)frame * planes + plane) * out_w * out_h;
    // any task tile in the (16 x (nby + 1)) tiles this segment reads?
    bool any = false;
    for (int k = threadIdx.x; k < 16 * (nby + 1); k += 256) {
        const int ux = bx0 - 1 + (k & 15), uy = by0 - 1 + (k >> 4);
        if (ux >= 0 && ux < g.tiles_x && uy >= 0 && uy < g.tiles_y && tile_map[(size_t)uy * g.tiles_x + ux] >= 0) any = true;
    }
    if (!__syncthreads_or(any)) {
        for (int k = threadIdx.x; k < nby * 8 * IMCLT_S * 8; k += 256) {
            const int row = by0 * 8 + k / (IMCLT_S * 8), x = bx0 * 8 + k % (IMCLT_S * 8);
            if (x < out_w) out[(size_t)row * out_w + x] = 0.f;
        }
        return;
    }
    float lp[8];
#pragma unroll
    for (int k = 0; k < 8; k++) lp[k] = 0.25f * (lpf ? lpf[kh * 8 + k] : 1.f);
    for (int i = threadIdx.x; i < 8 * IMCLT_S * 8; i += 256) (&s_carry[0][0])[i] = 0.f;
    for (int ty = by0 - 1; ty < by0 + nby; ty++) {
        bool valid = (ty >= 0) && (ty < g.tiles_y) && (tx >= 0) && (tx < g.tiles_x) &&
                     (tile_map[(size_t)(ty < 0 ? 0 : ty) * g.tiles_x + (tx < 0 ? 0 : tx)] >= 0);
        float p[8], q[8];
        if (valid) {
            const float* src = clt + (((size_t)cam * ntiles + (size_t)ty * g.tiles_x + tx) * g.colors + color) * 256 + kh * 8;
            float4 v0 = *reinterpret_cast<const float4*>(src + hs * 64);
            float4 v1 = *reinterpret_cast<const float4*>(src + hs * 64 + 4);
            float4 v2 = *reinterpret_cast<const float4*>(src + (hs | 2) * 64);
            float4 v3 = *reinterpret_cast<const float4*>(src + (hs | 2) * 64 + 4);
            p[0] = v0.x; p[1] = v0.y; p[2] = v0.z; p[3] = v0.w; p[4] = v1.x; p[5] = v1.y; p[6] = v1.z; p[7] = v1.w;
            q[0] = v2.x; q[1] = v2.y; q[2] = v2.z; q[3] = v2.w; q[4] = v3.x; q[5] = v3.y; q[6] = v3.z; q[7] = v3.w;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                p[k] *= lp[k];
                q[k] *= lp[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 8; k++) p[k] = q[k] = 0.f;
        }
        float row[16];
        imclt_unit(l, p, q, s_scr[unit], row);
#pragma unroll
        for (int c = 0; c < 16; c++) s_tile(unit, l, c) = row[c];
        __syncthreads();
        // assemble: output col (within strip) = s*8+c from tile s+1 (left half) and tile s (right half)
        if (threadIdx.x < IMCLT_S * 8) {
            const int colx = threadIdx.x, s = colx >> 3, c = colx & 7;
            const int x = (bx0 + s) * 8 + c;
            const bool wr = (ty >= by0) && (x < out_w);
#pragma unroll
            for (int rr = 0; rr < 8; rr++) {
                float top = s_tile(s + 1, rr, c) + s_tile(s, rr, 8 + c);
                float bot = s_tile(s + 1, rr + 8, c) + s_tile(s, rr + 8, 8 + c);
                if (wr) out[(size_t)(ty * 8 + rr) * out_w + x] = s_carry[rr][colx] + top;
                s_carry[rr][colx] = bot;
            }
        }
        __syncthreads();
    }
}

// ====================================================================================
// correlate2D: pairwise phase correlation (Correlation2d.java:197-389), combined strip
// (C2:556-790) and per-tile disparity (ImageDtt.java:2186-2238).
// One wavefront per task tile (4 tiles per block, no block barriers); lane = frequency bin in
// the transform-domain product, then lane-parallel DCT-IIe passes through 2 KiB of LDS.
// ====================================================================================
constexpr int CORR_MAX_STRIP_PAIRS = 8;

struct CorrArgs {
    int num_pairs;
    int cam1[CLT_MAX_PAIRS];
    int cam2[CLT_MAX_PAIRS];
    int strip_slot[CLT_MAX_PAIRS];             // index among the strip pairs or -1
    int num_strip;                             // pairs with a known orientation (<= 8)
    int strip_dir[CORR_MAX_STRIP_PAIRS];
    int strip_lg[CORR_MAX_STRIP_PAIRS];        // log2 of the interpolation denominator ss*(dir>1?1:2)
    int strip_is_hor1[CORR_MAX_STRIP_PAIRS];   // dir 0 and scale 1 (getMaskHorizontal(1))
    int strip_is_vert1[CORR_MAX_STRIP_PAIRS];  // dir 1 and scale 1 (getMaskVertical(1))
    int n_hor1, n_vert1;
    float col_w[3];
    float fat_zero, corr_offset, min_corr;
    int twice_diagonal, strip_height;
    float wndy[2], wndx[9], notch[15];
};

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// LDS floats per wavefront of k_corr2d
__host__ __device__ inline int corr_wave_floats(int num_strip) {
    return 256 + num_strip * CLT_CORR_LEN + 240 + 32 + 3;
}

// x^(1/n) for x > 0 with the hardware log2/exp2
__device__ __forceinline__ float pow_inv(float x, float inv_n) { return exp2f(log2f(x) * inv_n); }

// One sample of scaleRotateInterpoateSingleCorrelation (C2:610-718): the strip element (row,
// scol) of a pair with orientation dir, interpolation denominator 1 << lg.  NaN outside the range.
__device__ __forceinline__ float strip_value(const float* corr, int dir, int lg, int row, int scol) {
    const int mask = (1 << lg) - 1, ilimit = 7 << lg;
    const float rden = 1.f / (float)(1 << lg);
    float acc = 0.f;
    bool nan = false;
    for (int down = -1; down < ((row == 0) ? 0 : 2); down += 2) {
        int xnum, ynum;
        switch (dir) {
        case 0: xnum = 2 * scol + (row & 1); ynum = down * row; break;
        case 1: xnum = down * row; ynum = 2 * scol + (row & 1); break;
        case 2: xnum = scol - ((down * row) >> 1); ynum = scol + ((down * row + 1) >> 1); break;
        default: xnum = scol - ((down * row) >> 1); ynum = -scol - ((down * row + 1) >> 1); break;
        }
        if ((xnum > ilimit) || (xnum < -ilimit) || (ynum > ilimit) || (ynum < -ilimit)) {
            nan = true;   // the reference stores NaN there; a later += keeps it NaN
            continue;
        }
        const int ix0 = (xnum + ilimit) >> lg, iy0 = (ynum + ilimit) >> lg;
        const int fx = (xnum + ilimit) & mask, fy = (ynum + ilimit) & mask;   // exact dyadic fractions
        const float* c = corr + iy0 * 15 + ix0;
        float d;
        if (fx == 0 && fy == 0) {
            d = c[0];
        } else {
            const float dx = fx * rden, dy = fy * rden;
            if (fy == 0) d = (1.f - dx) * c[0] + dx * c[1];
            else if (fx == 0) d = (1.f - dy) * c[0] + dy * c[15];
            else d = ((1.f - dx) * (1.f - dy) * c[0]) + (dx * (1.f - dy) * c[1]) + ((1.f - dx) * dy * c[15]) + (dx * dy * c[16]);
        }
        acc = (row == 0) ? d : acc + 0.5f * d;
    }
    return nan ? __int_as_float(0x7fc00000) : acc;
}

// Default quad camera pairs (Correlation2d.java:52-58) for the specialised kernel
constexpr int QP_C1[6] = {0, 2, 0, 1, 0, 2};
constexpr int QP_C2[6] = {1, 3, 2, 3, 3, 1};
constexpr int QP_DIR[6] = {0, 0, 1, 1, 2, 3};
constexpr int QP_LG[6] = {1, 1, 1, 1, 0, 0};

// QUAD: 4 cameras, 3 colours, the 6 reference pairs -- the tile's CLT is held in registers and
// every pair/strip index is a compile-time constant; otherwise the pair table comes from CorrArgs.
// 2-D DCT-IIe / DST-IIe of the four quadrants of a pair's product (C2:316-319, DR:595-614) on the
// matrix cores, in place in the wavefront's 256 floats of LDS at S ([q][kh * 8 + kv] in, [q][y][x]
// out): the four quadrants as one 16x16 matrix Q[(qv, kv)][(qh, kh)]
// (q = 2 qv + qh, qv / qh: DST-IIe vertically / horizontally), out = By Q Bx^T with the
// block-diagonal By = Bx = diag(DCT-IIe, DST-IIe); two chains of four v_mfma_f32_16x16x4_f32.  Lane
// (g, n) = (lane >> 4, lane & 15): pass 1 (contraction over (qv, kv)) takes A = Q^T[(qh, kh) = n]
// [(qv, kv) = (g >> 1, 4 (g & 1) + s)] (one 16-byte LDS read) and B = u[s], the block-diagonal
// matrix at [row n][column 4g + s]; pass 2 (contraction over (qh, kh)) takes A = u[s] and B = pass
// 1's register s.  The result, (qh, x) = (g >> 1, 4 (g & 1) + r) and (qv, y) = (n >> 3, n & 7) in
// register r, goes back to S as [q][y][x] (one 16-byte write) for the unfold.
// QS: floats between the quadrants' 64-float blocks at S (16-byte aligned).
template <int QS = 64>
__device__ __forceinline__ void corr_dct_mx(float* S, const float (&u)[4], int lane) {
    const int g = lane >> 4, n = lane & 15;
    const float4 a = *reinterpret_cast<const float4*>(S + (2 * (g >> 1) + (n >> 3)) * QS + (n & 7) * 8 + 4 * (g & 1));
    const float av[4] = {a.x, a.y, a.z, a.w};
    float4x d1 = {0.f, 0.f, 0.f, 0.f}, d2 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; s++) d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s], u[s], d1, 0, 0, 0);
#pragma unroll
    for (int s = 0; s < 4; s++) d2 = __builtin_amdgcn_mfma_f32_16x16x4f32(u[s], d1[s], d2, 0, 0, 0);
    wave_sync();
    *reinterpret_cast<float4*>(S + (2 * (n >> 3) + (g >> 1)) * QS + (n & 7) * 8 + 4 * (g & 1)) =
        make_float4(d2[0], d2[1], d2[2], d2[3]);
    wave_sync();
}
// One pair's correlation from its frequency-domain product (one bin per lane, the four modes in
// acc): LPF, the 2-D DCT-IIe / DST-IIe above through the wavefront's 256 floats of LDS at S, unfold
// to the 15x15 tile (DttRad2.java:443-489): lane (ay, ax) = (lane >> 3, lane & 7) gets the values
// at (7 + ay, 7 + ax), (7 + ay, 7 - ax), (7 - ay, 7 + ax), (7 - ay, 7 - ax) in v[0..3].  The
// staging holds the four quadrants QS floats apart.
template <int QS = 64>
__device__ __forceinline__ void corr_fd_unfold(const float* acc, float lp, float* S, const float (&u)[4], int lane,
                                               float (&v)[4]) {
    const int y = lane >> 3, xx = lane & 7;
    wave_sync();
#pragma unroll
    for (int q = 0; q < 4; q++) S[q * QS + lane] = acc[q] * lp;
    wave_sync();
    corr_dct_mx<QS>(S, u, lane);
    // CC[ay][ax] +- SC[ay][ax-1] +- CS[ay-1][ax] +- SS[ay-1][ax-1]
    const float va = S[y * 8 + xx];
    const float vb = xx > 0 ? S[QS + y * 8 + xx - 1] : 0.f;
    const float vc = y > 0 ? S[2 * QS + (y - 1) * 8 + xx] : 0.f;
    const float vd = (xx > 0 && y > 0) ? S[3 * QS + (y - 1) * 8 + xx - 1] : 0.f;
    v[0] = 0.25f * (va + vb + vc + vd);
    v[1] = 0.25f * (va - vb + vc - vd);
    v[2] = 0.25f * (va + vb - vc - vd);
    v[3] = 0.25f * (va - vb - vc + vd);
}
// the unfolded values of corr_fd_unfold into the 15x15 tile at c (the lanes with ay or ax = 0 skip
// the duplicates)
__device__ __forceinline__ void corr_tile_put(float* c, const float (&v)[4], int lane) {
    const int y = lane >> 3, xx = lane & 7;
    c[(7 + y) * 15 + 7 + xx] = v[0];
    if (xx > 0) c[(7 + y) * 15 + 7 - xx] = v[1];
    if (y > 0) c[(7 - y) * 15 + 7 + xx] = v[2];
    if (xx > 0 && y > 0) c[(7 - y) * 15 + 7 - xx] = v[3];
}
// corr_fd_unfold written to cs (LDS) and, unless co is null, co (global)
__device__ __forceinline__ void corr_fd_to_tile_mx(const float* acc, float lp, float* S, const float (&u)[4], int lane,
                                                   float* co, float* cs) {
    float v[4];
    corr_fd_unfold(acc, lp, S, u, lane, v);
    corr_tile_put(cs, v, lane);
    if (co) corr_tile_put(co, v, lane);
}
// per-lane operand of corr_dct_mx: the block-diagonal DCT-IIe / DST-IIe at [n][4g + s]
__device__ __forceinline__ void corr_mx_operand(int lane, float (&u)[4]) {
    const int g = lane >> 4, n = lane & 15;
#pragma unroll
    for (int s = 0; s < 4; s++)
        u[s] = ((n >> 3) != (g >> 1)) ? 0.f : ((g >> 1) ? c_tab.siie[n & 7][4 * (g & 1) + s] : c_tab.ciie[n & 7][4 * (g & 1) + s]);
}

template <bool QUAD>
__global__ __launch_bounds__(256) void k_corr2d(const float* __restrict__ clt, const int32_t* __restrict__ tasks,
                                                int task_words, int num_tasks, const float* __restrict__ lpf,
                                                const CorrArgs* __restrict__ args_p, float* __restrict__ corr_out,
                                                float* __restrict__ combo_out, float* __restrict__ disp_out,
                                                const int32_t* __restrict__ task_frame, long long clt_fstride, GeomDev g) {
    extern __shared__ __attribute__((aligned(16))) float s_dyn[];
    const CorrArgs& A = *args_p;
    const int K = QUAD ? 3 : g.colors;
    const int P = QUAD ? 6 : A.num_pairs;
    const int NS = QUAD ? 6 : A.num_strip;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int t = xcd_block(blockIdx.x, gridDim.x) * 4 + wave;
    if (t >= num_tasks) return;                       // wave-uniform: no block barriers below
    const int wfl = (corr_wave_floats(NS) + 3) & ~3;
    float* S0 = s_dyn + wave * wfl;                   // [256]
    float* s_corr = S0 + 256;                         // [NS][225]
    float* s_combo = s_corr + NS * CLT_CORR_LEN;      // [240]
    float* s_c1d = s_combo + 240;                     // [2][16]
    const float qnan = __int_as_float(0x7fc00000);
    const int hw = A.strip_height, slen = hw * 15;
    const int32_t* rec = tasks + (size_t)t * task_words;
    if (rec[0] == 0) {
        for (int o = lane; o < P * CLT_CORR_LEN; o += 64) corr_out[(size_t)t * P * CLT_CORR_LEN + o] = qnan;
        if (NS) {
            if (combo_out)
                for (int e = lane; e < slen; e += 64) combo_out[(size_t)t * slen + e] = qnan;
            if (lane < CLT_DISP_LAYERS && disp_out)
                disp_out[(size_t)t * CLT_DISP_LAYERS + lane] =
                    (lane == CLT_DISP_STRENGTH || lane == CLT_DISP_HOR_STR || lane == CLT_DISP_VERT_STR ||
                     lane == CLT_DISP_CM_WEIGHT || lane == CLT_DISP_STR_RAW) ? 0.f : qnan;
        }
        return;
    }
    const int txy = rec[1];
    const size_t ntiles = (size_t)g.tiles_x * g.tiles_y;
    const size_t tile = (size_t)((txy >> 16) & 0xffff) * g.tiles_x + (txy & 0xffff);
    if (task_frame) clt += task_frame[t] * clt_fstride;   // batched frames
    // quad: the whole tile (4 cameras x 3 colours x 4 modes) for this lane's frequency bin
    float D[QUAD ? 4 : 1][3][4];
    if (QUAD) {
#pragma unroll
        for (int c = 0; c < (QUAD ? 4 : 1); c++)
#pragma unroll
            for (int col = 0; col < 3; col++)
#pragma unroll
                for (int n = 0; n < 4; n++) D[c][col][n] = clt[(((size_t)c * ntiles + tile) * 3 + col) * 256 + n * 64 + lane];
    }
    float umx[4];   // operand of the matrix-core DCT-IIe (corr_dct_mx)
    corr_mx_operand(lane, umx);
    const float lp = lpf[lane];
    float* corr_t = corr_out + (size_t)t * P * CLT_CORR_LEN;
#pragma unroll
    for (int p = 0; p < (QUAD ? 6 : 1); p++) {
        for (int pg = 0; pg < (QUAD ? 1 : P); pg++) {
            const int pp = QUAD ? p : pg;
            float acc[4] = {0.f, 0.f, 0.f, 0.f};
            const float* c1 = clt + ((size_t)A.cam1[pp] * ntiles + tile) * K * 256 + lane;
            const float* c2 = clt + ((size_t)A.cam2[pp] * ntiles + tile) * K * 256 + lane;
#pragma unroll
            for (int col = 0; col < (QUAD ? 3 : 1); col++) {
                for (int colg = 0; colg < (QUAD ? 1 : K); colg++) {
                    const int cc = QUAD ? col : colg;
                    float d1[4], d2[4];
#pragma unroll
                    for (int n = 0; n < 4; n++) {
                        if (QUAD) {
                            d1[n] = D[QUAD ? QP_C1[p] : 0][col][n];
                            d2[n] = D[QUAD ? QP_C2[p] : 0][col][n];
                        } else {
                            d1[n] = c1[cc * 256 + n * 64];
                            d2[n] = c2[cc * 256 + n * 64];
                        }
                    }
                    const float s1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2] + d1[3] * d1[3];
                    const float s2 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2] + d2[3] * d2[3];
                    const float a2 = sqrtf(s1 * s2);
                    float fz2 = 0.f;
                    if (A.fat_zero != 0.f) fz2 = wave_sum(a2) * (1.f / 64.f) * A.fat_zero * A.fat_zero;
                    const float scale = (1.f / (a2 + fz2)) * (K == 1 ? 1.f : A.col_w[cc]);
                    // product with the ZI sign/index table (Correlation2d.java:86-90)
                    const float t0 = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2] + d1[3] * d2[3];
                    const float t1 = -d1[1] * d2[0] + d1[0] * d2[1] - d1[3] * d2[2] + d1[2] * d2[3];
                    const float t2 = -d1[2] * d2[0] - d1[3] * d2[1] + d1[0] * d2[2] + d1[1] * d2[3];
                    const float t3 = d1[3] * d2[0] - d1[2] * d2[1] - d1[1] * d2[2] + d1[0] * d2[3];
                    acc[0] += t0 * scale;
                    acc[1] += t1 * scale;
                    acc[2] += t2 * scale;
                    acc[3] += t3 * scale;
                }
            }
#pragma unroll
            for (int n = 0; n < 4; n++) S0[n * 64 + lane] = acc[n] * lp;
            wave_sync();
            corr_dct_mx(S0, umx, lane);
            // unfold the 4 quadrants to 15x15 (DttRad2.java:443-489)
            const int slot = QUAD ? p : A.strip_slot[pp];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int o = lane + 64 * j;
                if (o < CLT_CORR_LEN) {
                    const int row = o / 15, cc = o - row * 15;
                    const int dy = row - 7, dx = cc - 7;
                    const int ay = dy < 0 ? -dy : dy, ax = dx < 0 ? -dx : dx;
                    float v = S0[ay * 8 + ax];
                    if (ax > 0) v += (dx > 0 ? 1.f : -1.f) * S0[64 + ay * 8 + ax - 1];
                    if (ay > 0) v += (dy > 0 ? 1.f : -1.f) * S0[128 + (ay - 1) * 8 + ax];
                    if (ax > 0 && ay > 0) v += ((dx > 0) == (dy > 0) ? 1.f : -1.f) * S0[192 + (ay - 1) * 8 + ax - 1];
                    v *= 0.25f;
                    corr_t[pp * CLT_CORR_LEN + o] = v;
                    if (slot >= 0) s_corr[slot * CLT_CORR_LEN + o] = v;
                }
            }
            wave_sync();
        }
    }
    if (NS == 0) return;
    // ---- combined strip: combineInterpolatedCorrelations of the rotated strips (C2:730-790)
    {
        const int ncomb = NS * (A.twice_diagonal ? 2 : 1);
        const float inv_n = 1.f / ncomb;
        for (int e = lane; e < slen; e += 64) {
            const int row = e / 15, scol = e - row * 15 - 7;
            float v;
            if (A.corr_offset >= 0.f) {
                v = 1.f;
#pragma unroll
                for (int sq = 0; sq < (QUAD ? 6 : 1); sq++) {
                    for (int sg = 0; sg < (QUAD ? 1 : NS); sg++) {
                        const int sidx = QUAD ? sq : sg;
                        const int dir = QUAD ? QP_DIR[sq] : A.strip_dir[sidx];
                        const int lg = QUAD ? QP_LG[sq] : A.strip_lg[sidx];
                        float d = strip_value(s_corr + sidx * CLT_CORR_LEN, dir, lg, row, scol) + A.corr_offset;
                        if (d < 0.f) d = 0.f;
                        v *= A.twice_diagonal ? d * d : d;
                    }
                }
                if (v > 0.f) v = pow_inv(v, inv_n) - A.corr_offset;
            } else {
                v = 0.f;
#pragma unroll
                for (int sq = 0; sq < (QUAD ? 6 : 1); sq++) {
                    for (int sg = 0; sg < (QUAD ? 1 : NS); sg++) {
                        const int sidx = QUAD ? sq : sg;
                        const int dir = QUAD ? QP_DIR[sq] : A.strip_dir[sidx];
                        const int lg = QUAD ? QP_LG[sq] : A.strip_lg[sidx];
                        v += (A.twice_diagonal ? 2.f : 1.f) * strip_value(s_corr + sidx * CLT_CORR_LEN, dir, lg, row, scol);
                    }
                }
                v *= inv_n;
            }
            s_combo[e] = v;
            if (combo_out) combo_out[(size_t)t * slen + e] = v;
        }
    }
    // ---- getMaxXSOrtho 1-D profiles (C2:2502-2651): lanes 0..14 horizontal, 32..46 vertical
    {
        const int which = lane >> 5, jj = lane & 31;
        const int npairs = which ? A.n_vert1 : A.n_hor1;
        if (jj < 15 && npairs > 0) {
            const float w = 1.f / npairs;
            float v = 0.f;
            for (int i = 0; i < 15; i++) {
                if (A.notch[i] == 0.f) continue;   // the combination is finite: adds exactly 0
                if (A.corr_offset >= 0.f) {
                    float cmb = 1.f;
                    for (int sidx = 0; sidx < NS; sidx++) {
                        if (!(which ? A.strip_is_vert1[sidx] : A.strip_is_hor1[sidx])) continue;
                        const float* c = s_corr + sidx * CLT_CORR_LEN;
                        const float d = which ? 0.5f * (c[jj * 15 + i] + c[jj * 15 + 14 - i])
                                              : 0.5f * (c[i * 15 + jj] + c[(14 - i) * 15 + jj]);
                        cmb *= d + A.corr_offset;
                    }
                    cmb = (cmb > 0.f) ? pow_inv(cmb, w) - A.corr_offset : -A.corr_offset;
                    v += cmb * A.notch[i];
                } else {
                    for (int sidx = 0; sidx < NS; sidx++) {
                        if (!(which ? A.strip_is_vert1[sidx] : A.strip_is_hor1[sidx])) continue;
                        const float* c = s_corr + sidx * CLT_CORR_LEN;
                        v += w * (which ? c[jj * 15 + i] : c[i * 15 + jj]) * A.notch[i];
                    }
                }
            }
            s_c1d[which * 16 + jj] = v;
        }
    }
    wave_sync();
    // ---- centre of mass on the strip axis, lane-parallel (getMaxXYInt C2:802-843, getMaxXCm C2:936-1032)
    int imx = 0;
    for (int i = 1; i < 15; i++)
        if (isnan(s_combo[imx]) || (s_combo[i] > s_combo[imx])) imx = i;
    const bool have_max = s_combo[imx] >= A.min_corr;
    const int ix = imx - 7;
    float s0 = 0.f, sx = 0.f, sx2 = 0.f;
    if (have_max) {
        // lanes 0..8: the 9 samples of row 0 (adx 0, then 2, 4, 6, 8 with dir -1 / +1), lanes
        // 16..23: the 8 samples of row 1 (adx 1, 3, 5, 7 with dir -1 / +1)
        const int dh = hw > 2 ? 2 : hw;
        const float wy_scale = (hw < 2) ? 1.f / A.wndy[0] : 1.f;
        const int dy = lane >> 4, k = lane & 15;
        int adx = -1, dir = 1;
        if (dy == 0 && k < 9) {
            adx = (k + 1) & ~1;
            dir = (k == 0 || (k & 1) == 0) ? 1 : -1;
        } else if (dy == 1 && k < 8) {
            adx = 2 * (k >> 1) + 1;
            dir = (k & 1) ? 1 : -1;
        }
        if (dy < dh && adx >= 0) {
            const int odd = dy & 1;
            const float wy = ((dy == 0) ? wy_scale : 2.f * wy_scale) * A.wndy[dy];
            const int idx = (adx * dir) >> 1;
            const int xq = 2 * idx + odd, x1 = 7 + ix + idx;
            if (x1 >= 0 && x1 < 15) {
                float d = s_combo[15 * dy + x1];
                if (!isnan(d) && d > 0.f) {
                    d *= wy * A.wndx[adx];
                    s0 = d;
                    sx = d * xq;
                    sx2 = d * xq * xq;
                }
            }
        }
        s0 = wave_sum(s0);
        sx = wave_sum(sx);
        sx2 = wave_sum(sx2);
    }
    if (lane == 0 && disp_out) {
        float res[CLT_DISP_LAYERS];
        for (int i = 0; i < CLT_DISP_LAYERS; i++) res[i] = qnan;
        res[CLT_DISP_STRENGTH] = res[CLT_DISP_HOR_STR] = res[CLT_DISP_VERT_STR] = res[CLT_DISP_CM_WEIGHT] = 0.f;
        res[CLT_DISP_STR_RAW] = 0.f;
        if (have_max) {
            res[CLT_DISP_INT] = (float)(-ix);
            res[CLT_DISP_STRENGTH] = s_combo[ix + 7];
            if (s0 != 0.f) {
                res[CLT_DISP_CM] = -(ix + sx / s0 / 2.f);
                res[CLT_DISP_CM_WEIGHT] = s0;
            }
        }
        for (int which = 0; which < 2; which++) {
            if ((which ? A.n_vert1 : A.n_hor1) == 0) continue;
            const float* v = s_c1d + which * 16;
            int ic = 0;
            for (int i = 1; i < 15; i++)
                if (v[i] > v[ic]) ic = i;
            float xc = (float)ic;
            if (ic > 0 && ic < 14) {   // 3-point parabola (PolynomialApproximation, degree 2)
                const float a = v[ic - 1], b = v[ic], c = v[ic + 1];
                const float den = a - 2.f * b + c;
                if (den != 0.f) xc = ic + (a - c) / (2.f * den);
            }
            const int icr = (int)floorf(xc + 0.5f);
            res[which ? CLT_DISP_VERT : CLT_DISP_HOR] = -(xc - 7.f);
            res[which ? CLT_DISP_VERT_STR : CLT_DISP_HOR_STR] = v[icr] / 8.f;
        }
        res[CLT_DISP_CM_RAW] = res[CLT_DISP_CM];   // kept for the LMA stage (k_corr_lma)
        res[CLT_DISP_STR_RAW] = res[CLT_DISP_STRENGTH];
        float* dout = disp_out + (size_t)t * CLT_DISP_LAYERS;
        for (int i = 0; i < CLT_DISP_LAYERS; i++) dout[i] = res[i];
    }
}

// ------------------------------------------------------------------------------------
// Quad camera form (4 cameras, 3 colours, the 6 pairs of Correlation2d.java:52-58): one wavefront
// per tile, the tile's 48 CLT values per frequency bin in registers; pair products pre-normalised
// per camera/colour; the 2-D DCT-IIe through 1 KiB of LDS; the 15x15 unfold by lane shuffles
// (lane (ay, ax) writes the up to 4 outputs (7 +- ay, 7 +- ax)); strips with tap addresses shared
// by the pairs of one orientation.
// ------------------------------------------------------------------------------------
// LDS per wavefront, a multiple of 4: the transform staging (256 floats, then the combined strip),
// the six correlation tiles and the two 1-D profiles
constexpr int CQ_WAVE_FLOATS = 256 + 6 * CLT_CORR_LEN + 32 + 2;

// One sample of a strip element (C2:610-718): position of (row, scol) for orientation DIR with
// interpolation denominator 1 << LG, sample side down = -1 / +1.  Returns false outside the
// correlation tile (the reference stores NaN); offsets of out-of-range or zero-weight taps point
// back at the base tap so every read stays inside the tile.
template <int DIR, int LG>
__device__ __forceinline__ bool strip_tap(int row, int scol, int down, int& base, int& ox, int& oy, float& fx, float& fy) {
    constexpr int ilimit = 7 << LG, mask = (1 << LG) - 1;
    int xnum, ynum;
    if (DIR == 0) { xnum = 2 * scol + (row & 1); ynum = down * row; }
    else if (DIR == 1) { xnum = down * row; ynum = 2 * scol + (row & 1); }
    else if (DIR == 2) { xnum = scol - ((down * row) >> 1); ynum = scol + ((down * row + 1) >> 1); }
    else { xnum = scol - ((down * row) >> 1); ynum = -scol - ((down * row + 1) >> 1); }
    const bool ok = (xnum <= ilimit) && (xnum >= -ilimit) && (ynum <= ilimit) && (ynum >= -ilimit);
    const int mx = (xnum + ilimit) & mask, my = (ynum + ilimit) & mask;
    base = ok ? ((ynum + ilimit) >> LG) * 15 + ((xnum + ilimit) >> LG) : 112;
    ox = (ok && mx) ? 1 : 0;
    oy = (ok && my) ? 15 : 0;
    fx = (float)mx * (1.f / (1 << LG));
    fy = (float)my * (1.f / (1 << LG));
    return ok;
}

// The taps of one strip element for one orientation: row 0 is one sample, other rows the mean
// of the samples at -row and +row (C2:640-700).
template <int DIR, int LG>
struct StripAt {
    int a, aox, aoy, b, box, boy;
    float wa[4], wb[4];
    bool nan;
    __device__ __forceinline__ StripAt(int row, int scol) {
        float fx, fy;
        const bool oka = strip_tap<DIR, LG>(row, scol, -1, a, aox, aoy, fx, fy);
        const float ka = (row == 0) ? 1.f : 0.5f;
        wa[0] = ka * (1.f - fx) * (1.f - fy);
        wa[1] = ka * fx * (1.f - fy);
        wa[2] = ka * (1.f - fx) * fy;
        wa[3] = ka * fx * fy;
        bool okb = true;
        if (row > 0) {
            okb = strip_tap<DIR, LG>(row, scol, 1, b, box, boy, fx, fy);
            wb[0] = 0.5f * (1.f - fx) * (1.f - fy);
            wb[1] = 0.5f * fx * (1.f - fy);
            wb[2] = 0.5f * (1.f - fx) * fy;
            wb[3] = 0.5f * fx * fy;
        } else {
            b = a; box = aox; boy = aoy;
            wb[0] = wb[1] = wb[2] = wb[3] = 0.f;
        }
        nan = !(oka && okb);
    }
    __device__ __forceinline__ float operator()(const float* c) const {
        if (LG == 0) return wa[0] * c[a] + wb[0] * c[b];
        return wa[0] * c[a] + wa[1] * c[a + aox] + wa[2] * c[a + aoy] + wa[3] * c[a + aox + aoy] +
               wb[0] * c[b] + wb[1] * c[b + box] + wb[2] * c[b + boy] + wb[3] * c[b + box + boy];
    }
};

// first maximum of 16 lanes (lane order), NaN as -inf; returns the index in every lane of the group
__device__ __forceinline__ int argmax16(float v, int idx) {
    if (isnan(v)) v = -INFINITY;
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
        const float v2 = __shfl_xor(v, o, 16);
        const int i2 = __shfl_xor(idx, o, 16);
        if (v2 > v || (v2 == v && i2 < idx)) { v = v2; idx = i2; }
    }
    return idx;
}


// Outputs of a quad task without correlation: NaN tiles and strip, zero strengths and weights.
__device__ __forceinline__ void corr_quad_empty(int t, int slen, int lane, float* __restrict__ corr_out,
                                                float* __restrict__ combo_out, float* __restrict__ disp_out) {
    const float qnan = __int_as_float(0x7fc00000);
    for (int o = lane; o < 6 * CLT_CORR_LEN; o += 64) corr_out[(size_t)t * 6 * CLT_CORR_LEN + o] = qnan;
    if (combo_out)
        for (int e = lane; e < slen; e += 64) combo_out[(size_t)t * slen + e] = qnan;
    if (lane < CLT_DISP_LAYERS && disp_out)
        disp_out[(size_t)t * CLT_DISP_LAYERS + lane] =
            (lane == CLT_DISP_STRENGTH || lane == CLT_DISP_HOR_STR || lane == CLT_DISP_VERT_STR ||
             lane == CLT_DISP_CM_WEIGHT || lane == CLT_DISP_STR_RAW) ? 0.f : qnan;
}

// The rest of correlate2D for a quad task (one wavefront) from its six tiles in LDS at s_corr: the
// combined strip (s_combo: LDS of 15 * strip_height floats, and combo_out), the two 1-D profiles
// (s_c1d: LDS of 32 floats) and the task's disparity layers.
__device__ __forceinline__ void corr_quad_tail(const CorrArgs& A, const float* s_corr, float* s_combo, float* s_c1d,
                                               int lane, int t, float* __restrict__ combo_out,
                                               float* __restrict__ disp_out) {
    const float qnan = __int_as_float(0x7fc00000);
    const int hw = A.strip_height, slen = hw * 15;
    // ---- combined strip (C2:730-790) of the 6 rotated strips (C2:556-718)
    {
        const int ncomb = 6 * (A.twice_diagonal ? 2 : 1);
        const float inv_n = 1.f / ncomb;
        for (int e = lane; e < slen; e += 64) {
            const int row = e / 15, scol = e - row * 15 - 7;
            const StripAt<0, 1> sh(row, scol);
            const StripAt<1, 1> sv(row, scol);
            const StripAt<2, 0> sd(row, scol);
            const StripAt<3, 0> sa(row, scol);
            float st[6] = {sh(s_corr), sh(s_corr + CLT_CORR_LEN), sv(s_corr + 2 * CLT_CORR_LEN), sv(s_corr + 3 * CLT_CORR_LEN),
                           sd(s_corr + 4 * CLT_CORR_LEN), sa(s_corr + 5 * CLT_CORR_LEN)};
            const bool nan = sh.nan || sv.nan || sd.nan || sa.nan;
            float v;
            if (A.corr_offset >= 0.f) {
                v = 1.f;
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    float dd = st[k] + A.corr_offset;
                    if (dd < 0.f) dd = 0.f;   // (keeps a NaN)
                    v *= A.twice_diagonal ? dd * dd : dd;
                }
                if (v > 0.f) v = fast_exp2(fast_log2(v) * inv_n) - A.corr_offset;
            } else {
                v = 0.f;
#pragma unroll
                for (int k = 0; k < 6; k++) v += st[k];
                v *= (A.twice_diagonal ? 2.f : 1.f) * inv_n;
            }
            if (nan) v = qnan;
            s_combo[e] = v;
            if (combo_out) combo_out[(size_t)t * slen + e] = v;
        }
    }
    // ---- getMaxXSOrtho 1-D profiles (C2:2502-2651) of the horizontal (pairs 0, 1: lanes 0..14)
    // and vertical (pairs 2, 3: lanes 32..46) scale-1 pairs; terms i and 14 - i are equal
    {
        const int which = lane >> 5, jj = lane & 31;
        if (jj < 15) {
            const float* ca = s_corr + (which ? 2 : 0) * CLT_CORR_LEN;
            const float* cb = ca + CLT_CORR_LEN;
            const int st_i = which ? 1 : 15, st_j = which ? 15 : 1;   // element (i, jj) offsets
            float v = 0.f;
            if (A.corr_offset >= 0.f) {
                for (int i = 0; i < 8; i++) {
                    const float nw = (i < 7) ? A.notch[i] + A.notch[14 - i] : A.notch[7];
                    if (nw == 0.f) continue;   // the combination is finite: adds exactly 0
                    const int o1 = i * st_i + jj * st_j, o2 = (14 - i) * st_i + jj * st_j;
                    const float da = 0.5f * (ca[o1] + ca[o2]) + A.corr_offset;
                    const float db = 0.5f * (cb[o1] + cb[o2]) + A.corr_offset;
                    const float cmb = da * db;
                    v += nw * ((cmb > 0.f) ? fast_sqrt(cmb) - A.corr_offset : -A.corr_offset);
                }
            } else {
                for (int i = 0; i < 15; i++) {
                    const int o1 = i * st_i + jj * st_j;
                    v += 0.5f * (ca[o1] + cb[o1]) * A.notch[i];
                }
            }
            s_c1d[which * 16 + jj] = v;
        }
    }
    wave_sync();
    // ---- integer maximum of the strip's middle row (getMaxXYInt, C2:802-843), centre of mass
    // (getMaxXCm, C2:936-1032) lane-parallel
    const int imx = __shfl(argmax16(lane < 15 ? s_combo[lane & 15] : -INFINITY, lane & 15), 0);
    const float vmax = s_combo[imx];
    const bool have_max = vmax >= A.min_corr;
    const int ix = imx - 7;
    float s0 = 0.f, sx = 0.f;
    if (have_max) {
        // lanes 0..8: the 9 samples of row 0 (adx 0, then 2, 4, 6, 8 with dir -1 / +1), lanes
        // 16..23: the 8 samples of row 1 (adx 1, 3, 5, 7 with dir -1 / +1)
        const int dh = hw > 2 ? 2 : hw;
        const float wy_scale = (hw < 2) ? 1.f / A.wndy[0] : 1.f;
        const int dy = lane >> 4, k = lane & 15;
        int adx = -1, dir = 1;
        if (dy == 0 && k < 9) {
            adx = (k + 1) & ~1;
            dir = (k == 0 || (k & 1) == 0) ? 1 : -1;
        } else if (dy == 1 && k < 8) {
            adx = 2 * (k >> 1) + 1;
            dir = (k & 1) ? 1 : -1;
        }
        if (dy < dh && adx >= 0) {
            const int odd = dy & 1;
            const float wy = ((dy == 0) ? wy_scale : 2.f * wy_scale) * A.wndy[dy];
            const int idx = (adx * dir) >> 1;
            const int xq = 2 * idx + odd, x1 = 7 + ix + idx;
            if (x1 >= 0 && x1 < 15) {
                float dv = s_combo[15 * dy + x1];
                if (!isnan(dv) && dv > 0.f) {
                    dv *= wy * A.wndx[adx];
                    s0 = dv;
                    sx = dv * xq;
                }
            }
        }
        s0 = wave_sum(s0);
        sx = wave_sum(sx);
    }
    // ---- maxima of the ortho profiles with the 3-point parabola (PolynomialApproximation, degree 2)
    float xo = 0.f, so = 0.f;
    {
        const int which = lane >> 5, jj = lane & 15;
        const float* v = s_c1d + which * 16;
        const int ic = argmax16((lane & 31) < 15 ? v[jj] : -INFINITY, jj);
        float xc = (float)ic;
        if (ic > 0 && ic < 14) {
            const float a = v[ic - 1], b = v[ic], c = v[ic + 1];
            const float den = a - 2.f * b + c;
            if (den != 0.f) xc = ic + (a - c) / (2.f * den);
        }
        const int icr = (int)floorf(xc + 0.5f);
        xo = -(xc - 7.f);
        so = v[icr] * 0.125f;
    }
    const float hor = __shfl(xo, 0), hor_s = __shfl(so, 0), vert = __shfl(xo, 32), vert_s = __shfl(so, 32);
    if (lane < CLT_DISP_LAYERS && disp_out) {
        float r = qnan;
        switch (lane) {
        case CLT_DISP_INT: r = have_max ? (float)(-ix) : qnan; break;
        case CLT_DISP_CM:
        case CLT_DISP_CM_RAW: r = (have_max && s0 != 0.f) ? -(ix + sx / s0 / 2.f) : qnan; break;
        case CLT_DISP_STRENGTH:
        case CLT_DISP_STR_RAW: r = have_max ? vmax : 0.f; break;
        case CLT_DISP_POLY:
        case CLT_DISP_POLY_STR: r = qnan; break;
        case CLT_DISP_HOR: r = hor; break;
        case CLT_DISP_HOR_STR: r = hor_s; break;
        case CLT_DISP_VERT: r = vert; break;
        case CLT_DISP_VERT_STR: r = vert_s; break;
        default: r = have_max ? s0 : 0.f; break;   // CLT_DISP_CM_WEIGHT
        }
        disp_out[(size_t)t * CLT_DISP_LAYERS + lane] = r;
    }
}

template <bool FZ>   // FZ: fat zero in use (the per-bin normaliser needs the mean over the bins)
__global__ __launch_bounds__(256, FZ ? 4 : 6) void k_corr2d_quad(const float* __restrict__ clt, const int32_t* __restrict__ tasks,
                                                     int task_words, int num_tasks, const float* __restrict__ lpf,
                                                     const CorrArgs* __restrict__ args_p, float* __restrict__ corr_out,
                                                     float* __restrict__ combo_out, float* __restrict__ disp_out,
                                                     const int32_t* __restrict__ task_frame, long long clt_fstride,
                                                     GeomDev g, float* __restrict__ mix_out) {
    extern __shared__ __attribute__((aligned(16))) float s_dyn[];
    const CorrArgs& A = *args_p;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int t = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + wave;   // one tile per wavefront
    if (t >= num_tasks) return;                       // wave-uniform: no block barriers below
    float* S = s_dyn + wave * CQ_WAVE_FLOATS;         // [4][64] transform staging
    float* s_corr = S + 256;                          // [6][225]
    float* s_combo = S;                               // [240], once the pairs are done with S
    float* s_c1d = s_corr + 6 * CLT_CORR_LEN;         // [2][16]
    const int32_t* rec = tasks + (size_t)t * task_words;
    if (rec[0] == 0) {
        corr_quad_empty(t, A.strip_height * 15, lane, corr_out, combo_out, disp_out);
        return;
    }
    const int txy = rec[1];
    const size_t ntiles = (size_t)g.tiles_x * g.tiles_y;
    const size_t tile = (size_t)((txy >> 16) & 0xffff) * g.tiles_x + (txy & 0xffff);
    if (task_frame) clt += task_frame[t] * clt_fstride;   // batched frames
    float D[4][3][4];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int col = 0; col < 3; col++)
#pragma unroll
            for (int n = 0; n < 4; n++) D[c][col][n] = clt[(((size_t)c * ntiles + tile) * 3 + col) * 256 + n * 64 + lane];
    if (mix_out) {   // cltMixCameras (C2:425-447): the mean of the 4 cameras, for the rig's inter-camera pair
#pragma unroll
        for (int col = 0; col < 3; col++)
#pragma unroll
            for (int n = 0; n < 4; n++)
                mix_out[((size_t)t * 3 + col) * 256 + n * 64 + lane] =
                    0.25f * (D[0][col][n] + D[1][col][n] + D[2][col][n] + D[3][col][n]);
    }
    constexpr bool fz = FZ;
    float nrm[4][3];   // per bin sum over the 4 modes of the squares (C2:205-212)
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int col = 0; col < 3; col++)
            nrm[c][col] = D[c][col][0] * D[c][col][0] + D[c][col][1] * D[c][col][1] + D[c][col][2] * D[c][col][2] +
                          D[c][col][3] * D[c][col][3];
    if (!fz) {   // 1 / sqrt(n1 n2) * w = (sqrt(w) / sqrt(n1)) * (sqrt(w) / sqrt(n2)): scale each camera once
#pragma unroll
        for (int col = 0; col < 3; col++) {
            const float sw = fast_sqrt(A.col_w[col]);
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const float k = sw * __builtin_amdgcn_rsqf(nrm[c][col]);
#pragma unroll
                for (int n = 0; n < 4; n++) D[c][col][n] *= k;
            }
        }
    }
    float umx[4];
    corr_mx_operand(lane, umx);
    const float lp = lpf[lane];
    float* corr_t = corr_out + (size_t)t * 6 * CLT_CORR_LEN;
#pragma unroll
    for (int p = 0; p < 6; p++) {
        const int c1 = QP_C1[p], c2 = QP_C2[p];
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int col = 0; col < 3; col++) {
            const float* d1 = D[c1][col];
            const float* d2 = D[c2][col];
            float scale = 1.f;
            if (fz) {   // fat zero: mean over the bins of sqrt(n1 n2) (C2:213-218)
                const float a2 = fast_sqrt(nrm[c1][col] * nrm[c2][col]);
                const float fz2 = wave_sum(a2) * (1.f / 64.f) * A.fat_zero * A.fat_zero;
                scale = fast_rcp(a2 + fz2) * A.col_w[col];
            }
            // product with the ZI sign/index table (Correlation2d.java:86-90)
            const float t0 = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2] + d1[3] * d2[3];
            const float t1 = -d1[1] * d2[0] + d1[0] * d2[1] - d1[3] * d2[2] + d1[2] * d2[3];
            const float t2 = -d1[2] * d2[0] - d1[3] * d2[1] + d1[0] * d2[2] + d1[1] * d2[3];
            const float t3 = d1[3] * d2[0] - d1[2] * d2[1] - d1[1] * d2[2] + d1[0] * d2[3];
            acc[0] = fmaf(t0, scale, acc[0]);
            acc[1] = fmaf(t1, scale, acc[1]);
            acc[2] = fmaf(t2, scale, acc[2]);
            acc[3] = fmaf(t3, scale, acc[3]);
        }
        corr_fd_to_tile_mx(acc, lp, S, umx, lane, nullptr, s_corr + p * CLT_CORR_LEN);
    }
    wave_sync();
    // the six correlation tiles, contiguous in LDS and in the output: stored as 8-byte pairs, each
    // store instruction a contiguous 512-byte run (the tile's base is 8-byte aligned)
    for (int o = lane; o < 3 * CLT_CORR_LEN; o += 64)
        reinterpret_cast<float2*>(corr_t)[o] = reinterpret_cast<const float2*>(s_corr)[o];
    corr_quad_tail(A, s_corr, s_combo, s_c1d, lane, t, combo_out, disp_out);
}

// convert_direct and correlate2D of one quad task (4 cameras, 3 colours, the 6 pairs) per
// 3-wavefront block, the CLT passed from one to the other in LDS: after the conversion (the CLT
// also goes to memory for the textures) wavefront w correlates pairs 2 w and 2 w + 1 from the
// block's 12 (camera, colour) units in LDS, then the six tiles meet in LDS for wavefront 0's strip,
// profiles and disparity layers while wavefronts 1 and 2 store the tiles.  Same per-pair arithmetic
// as k_corr2d_quad (pair products of normalised vectors, the matrix-core DCT-IIe, the unfold).  The
// transform staging of wavefront w is in the 80 unused floats after each of its units' 256
// coefficients (one quadrant per unit), so the block needs no LDS beyond the conversion's; at 72
// VGPRs 7 wavefronts per SIMD.
template <bool FZ>
__global__ __launch_bounds__(192, 7) void k_convert_corr(const float* __restrict__ images, const float* __restrict__ kernels,
                                                      const float* __restrict__ koffsets, const int32_t* __restrict__ tasks,
                                                      int task_words, int num_tasks, float* __restrict__ clt,
                                                      const float* __restrict__ lpf, int lpf_mask, double inv_kstep,
                                                      const int32_t* __restrict__ task_frame, long long img_fstride,
                                                      long long clt_fstride, GeomDev g, const float* __restrict__ lpf_corr,
                                                      const CorrArgs* __restrict__ args_p, float* __restrict__ corr_out,
                                                      float* __restrict__ combo_out, float* __restrict__ disp_out) {
    static_assert(CONV_UNIT >= 256 + 64 && CONV_UNIT % 4 == 0, "a quadrant in each unit's tail");
    __shared__ __attribute__((aligned(16))) float s_lds[12 * CONV_UNIT];
    float (*units)[CONV_UNIT] = reinterpret_cast<float (*)[CONV_UNIT]>(s_lds);
    int t;
    // one task per block: every wavefront gets the same answer, so the return is block-uniform
    const bool busy = convert_tiles(images, kernels, koffsets, tasks, task_words, num_tasks, clt, lpf, lpf_mask,
                                    inv_kstep, task_frame, img_fstride, clt_fstride, g, units, t);
    const CorrArgs& A = *args_p;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (!busy) {
        if (wave == 0) corr_quad_empty(t, A.strip_height * 15, lane, corr_out, combo_out, disp_out);
        return;
    }
    if (!FZ) {   // each unit's bins pre-normalised once, sqrt(w) / |D| (C2:205-212), in place (the
                 // lane rescales the 16 floats it just stored from, so no other lane is involved)
        float* u = units[threadIdx.x >> 4] + 4 * (threadIdx.x & 15);
        const float sw = fast_sqrt(A.col_w[wave]);   // wavefront w converted colour w
        float4 m[4];
#pragma unroll
        for (int q = 0; q < 4; q++) m[q] = *reinterpret_cast<const float4*>(u + q * 64);
        const float k0 = sw * __builtin_amdgcn_rsqf(m[0].x * m[0].x + m[1].x * m[1].x + m[2].x * m[2].x + m[3].x * m[3].x);
        const float k1 = sw * __builtin_amdgcn_rsqf(m[0].y * m[0].y + m[1].y * m[1].y + m[2].y * m[2].y + m[3].y * m[3].y);
        const float k2 = sw * __builtin_amdgcn_rsqf(m[0].z * m[0].z + m[1].z * m[1].z + m[2].z * m[2].z + m[3].z * m[3].z);
        const float k3 = sw * __builtin_amdgcn_rsqf(m[0].w * m[0].w + m[1].w * m[1].w + m[2].w * m[2].w + m[3].w * m[3].w);
#pragma unroll
        for (int q = 0; q < 4; q++)
            *reinterpret_cast<float4*>(u + q * 64) = make_float4(m[q].x * k0, m[q].y * k1, m[q].z * k2, m[q].w * k3);
    }
    __syncthreads();   // the 12 units' CLT in LDS
    float* S = units[wave * 4] + 256;   // quadrant q in the tail of unit 4 wave + q
    float umx[4];
    corr_mx_operand(lane, umx);
    const float lp = lpf_corr[lane];
    float v[2][4];
#pragma unroll
    for (int p = 0; p < 6; p++) {
        if ((p >> 1) != wave) continue;   // wave-uniform
        const int c1 = QP_C1[p], c2 = QP_C2[p];
        float d1[3][4], d2[3][4], n1[3], n2[3];   // (without fat zero: already normalised)
#pragma unroll
        for (int col = 0; col < 3; col++) {
#pragma unroll
            for (int n = 0; n < 4; n++) {
                d1[col][n] = units[col * 4 + c1][n * 64 + lane];
                d2[col][n] = units[col * 4 + c2][n * 64 + lane];
            }
            if (FZ) {
                n1[col] = d1[col][0] * d1[col][0] + d1[col][1] * d1[col][1] + d1[col][2] * d1[col][2] + d1[col][3] * d1[col][3];
                n2[col] = d2[col][0] * d2[col][0] + d2[col][1] * d2[col][1] + d2[col][2] * d2[col][2] + d2[col][3] * d2[col][3];
            }
        }
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int col = 0; col < 3; col++) {
            const float* a = d1[col];
            const float* b = d2[col];
            float scale = 1.f;
            if (FZ) {   // fat zero: mean over the bins of sqrt(n1 n2) (C2:213-218)
                const float a2 = fast_sqrt(n1[col] * n2[col]);
                const float fz2 = wave_sum(a2) * (1.f / 64.f) * A.fat_zero * A.fat_zero;
                scale = fast_rcp(a2 + fz2) * A.col_w[col];
            }
            // product with the ZI sign/index table (Correlation2d.java:86-90)
            const float t0 = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
            const float t1 = -a[1] * b[0] + a[0] * b[1] - a[3] * b[2] + a[2] * b[3];
            const float t2 = -a[2] * b[0] - a[3] * b[1] + a[0] * b[2] + a[1] * b[3];
            const float t3 = a[3] * b[0] - a[2] * b[1] - a[1] * b[2] + a[0] * b[3];
            acc[0] = fmaf(t0, scale, acc[0]);
            acc[1] = fmaf(t1, scale, acc[1]);
            acc[2] = fmaf(t2, scale, acc[2]);
            acc[3] = fmaf(t3, scale, acc[3]);
        }
        corr_fd_unfold<CONV_UNIT>(acc, lp, S, umx, lane, v[p & 1]);
    }
    __syncthreads();   // every wavefront is done with the units: the tiles go where they were
    float* s_corr = s_lds;                    // [6][225]
    float* s_c1d = s_lds + 6 * CLT_CORR_LEN + 2;   // [2][16]
    float* s_combo = s_c1d + 32;                    // [15 * strip_height <= 240]
    corr_tile_put(s_corr + (2 * wave) * CLT_CORR_LEN, v[0], lane);
    corr_tile_put(s_corr + (2 * wave + 1) * CLT_CORR_LEN, v[1], lane);
    __syncthreads();
    if (wave) {   // the six tiles, contiguous in LDS and in the output, as 8-byte pairs
        float* corr_t = corr_out + (size_t)t * 6 * CLT_CORR_LEN;
        for (int o = threadIdx.x - 64; o < 3 * CLT_CORR_LEN; o += 128)
            reinterpret_cast<float2*>(corr_t)[o] = reinterpret_cast<const float2*>(s_corr)[o];
    } else {
        corr_quad_tail(A, s_corr, s_combo, s_c1d, lane, t, combo_out, disp_out);
    }
}

// ====================================================================================
// Correlation-peak loss and its adjoint (quad camera, 3 colours, the 6 reference pairs).
//
// Per task t the loss is l_t = -sum_p C_p, C_p the zero-offset value of pair p's phase correlation,
// i.e. the centre (7, 7) of its 15x15 tile.  With the DCT-IIe and unfold of correlate2D this is
//   C_p = 1/16 sum_colours w_k sum_bins lpf_bin <D1, D2>_bin / (|D1|_bin |D2|_bin)   (fat zero 0)
// over the 64 frequency bins of the 4-mode CLT vectors D of the two cameras.  One workgroup per
// task does convert_direct for its 4 cameras x 3 colours together with the tangents of the CLT by
// the residual shifts (rx, ry) of every unit (forward mode through the shifted-window fold, the
// transforms, the kernel multiply and the phase rotation), then the reverse pass of the loss to
// dl/dD in LDS, and contracts it with the tangents: dl/d(rx, ry) per unit, dl/d(x, y) of every port
// centre through the kernel-offset chain of the window position (ImageDtt.java:4915-4943).
// Also returned per pair: the curvatures of C_p under a relative shift along x and y
// (-d2C/dtau2 = 1/16 sum w lpf omega^2 f, omega the phase rate of the bin), used as the
// Gauss-Newton weights of the LMA.
// Output per task (CORR_ADJ_STRIDE floats): loss, C[6], kx[6], ky[6], dl/dxy[4][2].
// ====================================================================================
constexpr int CORR_ADJ_STRIDE = 32;

__global__ __launch_bounds__(192) void k_corr_adjoint_quad(const float* __restrict__ images, const float* __restrict__ kernels,
                                                           const float* __restrict__ koffsets, const int32_t* __restrict__ tasks,
                                                           int task_words, int num_tasks, const float* __restrict__ lpf,
                                                           const CorrArgs* __restrict__ args_p, double inv_kstep,
                                                           float* __restrict__ out, GeomDev g) {
    __shared__ __attribute__((aligned(16))) float s_buf[24][CONV_UNIT];   // unit u: value rows u, tangent rows 12 + u
    __shared__ float s_red[3][20];
    __shared__ float s_gxy[4][3][2];
    const CorrArgs& A = *args_p;
    const int t = xcd_block(blockIdx.x, gridDim.x);
    if (t >= num_tasks) return;                              // block-uniform
    const int color = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave = colour
    const int unit_in_block = threadIdx.x >> 4, l = threadIdx.x & 15;
    const int cam = (threadIdx.x >> 4) & 3;
    const int32_t* rec = tasks + (size_t)t * task_words;
    float* o = out + (size_t)t * CORR_ADJ_STRIDE;
    if (rec[0] == 0) {
        if (threadIdx.x < CORR_ADJ_STRIDE) o[threadIdx.x] = (threadIdx.x == 0) ? __int_as_float(0x7fc00000) : 0.f;
        return;
    }
    const int txy = rec[1];
    const double cx = (double)__int_as_float(rec[2 + 2 * cam]);
    const double cy = (double)__int_as_float(rec[3 + 2 * cam]);
    const float* ce = nullptr;
    const float* kt = nullptr;
    int ktx = 0, kty = 0;
    if (kernels) {
        ktx = (int)floor(cx * inv_kstep + 0.5) + 1;
        kty = (int)floor(cy * inv_kstep + 0.5) + 1;
        kty = kty < 0 ? 0 : (kty >= g.kv ? g.kv - 1 : kty);
        ktx = ktx < 0 ? 0 : (ktx >= g.kh ? g.kh - 1 : ktx);
        const size_t kidx = (((size_t)cam * g.kv + kty) * g.kh + ktx) * 3 + color;
        ce = koffsets + kidx * 8;
        kt = kernels + kidx * 256;
    }
    const TilePos pos = tile_position(cx, cy, ce, g.kstep, ktx, kty);
    const int left = pos.left, top = pos.top;
    const float rx = pos.rx, ry = pos.ry;
    // a task with any window outside the frame is left out (NaN loss, zero gradient): clamped
    // border windows are nearly constant and would correlate perfectly
    const bool inside = (left >= 0) && (left + 16 <= g.width) && (top >= 0) && (top + 16 <= g.height);
    if (__syncthreads_or(!inside)) {
        if (threadIdx.x < CORR_ADJ_STRIDE) o[threadIdx.x] = (threadIdx.x == 0) ? __int_as_float(0x7fc00000) : 0.f;
        return;
    }
    float* sb = s_buf[unit_in_block];
    float* sbx = s_buf[12 + unit_in_block];
    {   // window rows, coalesced (lane = column), this colour only, parity-preserving clamp
        const float* img = images + (size_t)cam * g.height * g.width;
        const int pc = left + l, cp = pc & 1;
        auto colour_at = [&](int rp) {
            const bool rr = rp == g.red_row, rc = cp == g.red_col;
            return rr ? (rc ? 0 : 2) : (rc ? 2 : 1);
        };
        const bool keep_e = colour_at(top & 1) == color, keep_o = colour_at((top & 1) ^ 1) == color;
        const int pcc = clamp_parity(pc, g.width);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int pr = clamp_parity(top + r, g.height);
            sb[r * CONV_ROW + l] = ((r & 1) ? keep_o : keep_e) ? img[(size_t)pr * g.width + pcc] : 0.f;
        }
    }
    wave_sync();
    const int r = l;
    const bool row_red = (((top + r) & 1) == g.red_row);
    const int cpar = (color == 0) ? (row_red ? g.red_col : -1) : (color == 1) ? (row_red ? -1 : (g.red_col ^ 1))
                                                                               : (row_red ? (g.red_col ^ 1) : g.red_col);
    float x[16];
#pragma unroll
    for (int j = 0; j < 16; j += 4) {
        const float4 v = *reinterpret_cast<const float4*>(&sb[r * CONV_ROW + j]);
        x[j] = v.x; x[j + 1] = v.y; x[j + 2] = v.z; x[j + 3] = v.w;
    }
    wave_sync();
    constexpr float kPi16 = 3.14159265358979f / 16.f;
    float ahs, ahc, avs, avc;
    sincos_pi(rx * 0.0625f, &ahs, &ahc);
    sincos_pi(ry * 0.0625f, &avs, &avc);
    {   // stage H: value with (ahc, ahs), tangent by rx with (d ahc, d ahs) = pi/16 (-ahs, ahc)
        float yc[8], ys[8], hc[8], hsn[8];
        const float sg = (((cpar ^ left) & 1) ? -1.f : 1.f);
        fold16(x, ahc, ahs, yc, ys);
        dct_iv8(yc, hc);
#pragma unroll
        for (int k = 0; k < 8; k++) hsn[k] = sg * hc[7 - k];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            sb[k * CONV_ROW + r] = hc[k];
            sb[(8 + k) * CONV_ROW + r] = hsn[k];
        }
        fold16(x, -ahs * kPi16, ahc * kPi16, yc, ys);
        dct_iv8(yc, hc);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            sbx[k * CONV_ROW + r] = hc[k];
            sbx[(8 + k) * CONV_ROW + r] = sg * hc[7 - k];
        }
    }
    wave_sync();
    const int hs = l >> 3, kh = l & 7;
    float av[8], bv[8], ax[8], bx[8], ay[8], by[8];
    {   // stage V: value V(col), by rx V(col_x), by ry V'(col)
        float col[16], colx[16];
#pragma unroll
        for (int i = 0; i < 16; i += 4) {
            const float4 v = *reinterpret_cast<const float4*>(&sb[(hs * 8 + kh) * CONV_ROW + i]);
            const float4 w = *reinterpret_cast<const float4*>(&sbx[(hs * 8 + kh) * CONV_ROW + i]);
            col[i] = v.x; col[i + 1] = v.y; col[i + 2] = v.z; col[i + 3] = v.w;
            colx[i] = w.x; colx[i + 1] = w.y; colx[i + 2] = w.z; colx[i + 3] = w.w;
        }
        const bool rb = color != 2;
        const int rpar = (color == 0) ? g.red_row : (g.red_row ^ 1);
        const float sgv = (((rpar ^ top) & 1) ? -1.f : 1.f);
        auto vstage = [&](const float* c, float vc, float vs, float* oa, float* ob) {
            float zc[8], zs[8];
            fold16(c, vc, vs, zc, zs);
            dct_iv8(zc, oa);
            if (rb) {
#pragma unroll
                for (int k = 0; k < 8; k++) ob[k] = sgv * oa[7 - k];
            } else {
                dst_iv8(zs, ob);
            }
        };
        vstage(col, avc, avs, av, bv);
        vstage(colx, avc, avs, ax, bx);
        vstage(col, -avs * kPi16, avc * kPi16, ay, by);
    }
    const float sigma = hs ? 1.f : -1.f;
    if (kt) {   // kernel multiply (linear: same for value and tangents)
        float k0[8], k1[8], k2[8], k3[8];
        const float* kq = kt + kh * 8;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            k0[k] = kq[k];
            k1[k] = kq[64 + k];
            k2[k] = kq[128 + k];
            k3[k] = kq[192 + k];
        }
        auto kmul = [&](float* a, float* b) {
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const float pa = lane_xor8(a[k]) * sigma, pb = lane_xor8(b[k]) * sigma;
                const float oa = a[k] * k0[k] + pa * k1[k] - b[k] * k2[k] - pb * k3[k];
                const float ob = b[k] * k0[k] + pb * k1[k] + a[k] * k2[k] + pa * k3[k];
                a[k] = oa;
                b[k] = ob;
            }
        };
        kmul(av, bv);
        kmul(ax, bx);
        kmul(ay, by);
    }
    {   // phase rotation: value, and the product rule for both tangents
        float sh, ch, cv[8], sv[8];
        sincos_pi((kh + 0.5f) * rx * 0.125f, &sh, &ch);
        half_phase8(ry * 0.125f, cv, sv);
        const float wh = 3.14159265358979f * (kh + 0.5f) * 0.125f;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float wv = 3.14159265358979f * (k + 0.5f) * 0.125f;
            const float pa = lane_xor8(av[k]) * sigma, pb = lane_xor8(bv[k]) * sigma;
            const float pax = lane_xor8(ax[k]) * sigma, pbx = lane_xor8(bx[k]) * sigma;
            const float pay = lane_xor8(ay[k]) * sigma, pby = lane_xor8(by[k]) * sigma;
            const float ta = av[k] * ch + pa * sh, tb = bv[k] * ch + pb * sh;                 // S_h Y
            const float tax = ax[k] * ch + pax * sh + wh * (pa * ch - av[k] * sh);            // d/drx S_h Y
            const float tbx = bx[k] * ch + pbx * sh + wh * (pb * ch - bv[k] * sh);
            const float tay = ay[k] * ch + pay * sh, tby = by[k] * ch + pby * sh;             // S_h dY/dry
            av[k] = ta * cv[k] - tb * sv[k];
            bv[k] = tb * cv[k] + ta * sv[k];
            ax[k] = tax * cv[k] - tbx * sv[k];
            bx[k] = tbx * cv[k] + tax * sv[k];
            ay[k] = tay * cv[k] - tby * sv[k] - wv * (ta * sv[k] + tb * cv[k]);
            by[k] = tby * cv[k] + tay * sv[k] + wv * (ta * cv[k] - tb * sv[k]);
        }
    }
    __syncthreads();   // every wavefront is done with its staging rows
    float* V = &s_buf[0][0];   // [cam][colour][mode][kh][kv], then dl/dCLT in place
    {
        float* vu = V + (cam * 3 + color) * 256 + kh * 8;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            vu[hs * 64 + k] = av[k];
            vu[(hs | 2) * 64 + k] = bv[k];
        }
    }
    __syncthreads();
    {   // reverse pass of the loss: thread = (colour, bin)
        const int k = threadIdx.x >> 6, bin = threadIdx.x & 63;
        const int bh = bin >> 3, bvv = bin & 7;   // bin = kh * 8 + kv
        float D[4][4], rr[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            float n2 = 0.f;
#pragma unroll
            for (int n = 0; n < 4; n++) {
                D[c][n] = V[(c * 3 + k) * 256 + n * 64 + bin];
                n2 += D[c][n] * D[c][n];
            }
            rr[c] = __builtin_amdgcn_rsqf(n2);
        }
        const float lw = 0.0625f * lpf[bin] * A.col_w[k];
        const float oh = 3.14159265358979f * (bh + 0.5f) * 0.125f, ov = 3.14159265358979f * (bvv + 0.5f) * 0.125f;
        float G[4][4] = {};
        float part[18];
#pragma unroll
        for (int p = 0; p < 6; p++) {
            const int c1 = QP_C1[p], c2 = QP_C2[p];
            float dot = 0.f;
#pragma unroll
            for (int n = 0; n < 4; n++) dot += D[c1][n] * D[c2][n];
            const float r12 = rr[c1] * rr[c2], f = dot * r12;
            part[p] = lw * f;
            part[6 + p] = lw * oh * oh * f;
            part[12 + p] = lw * ov * ov * f;
#pragma unroll
            for (int n = 0; n < 4; n++) {   // d(-C_p)/dD
                G[c1][n] -= lw * (D[c2][n] * r12 - f * D[c1][n] * rr[c1] * rr[c1]);
                G[c2][n] -= lw * (D[c1][n] * r12 - f * D[c2][n] * rr[c2] * rr[c2]);
            }
        }
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int n = 0; n < 4; n++) V[(c * 3 + k) * 256 + n * 64 + bin] = G[c][n];
#pragma unroll
        for (int q = 0; q < 18; q++) {
            const float s = wave_sum(part[q]);
            if ((threadIdx.x & 63) == 0) s_red[k][q] = s;
        }
    }
    __syncthreads();
    {   // contraction with the tangents over this unit's 256 coefficients
        const float* gu = V + (cam * 3 + color) * 256 + kh * 8;
        float grx = 0.f, gry = 0.f;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float ga = gu[hs * 64 + k], gb = gu[(hs | 2) * 64 + k];
            grx += ga * ax[k] + gb * bx[k];
            gry += ga * ay[k] + gb * by[k];
        }
#pragma unroll
        for (int ofs = 8; ofs > 0; ofs >>= 1) {
            grx += __shfl_xor(grx, ofs, 16);
            gry += __shfl_xor(gry, ofs, 16);
        }
        if (l == 0) {   // rx = left - px, px = cx - 8 - (c0 + c4 kdx + c5 kdy), py likewise (ImageDtt.java:4915-4943)
            const float e4 = ce ? ce[4] : 0.f, e5 = ce ? ce[5] : 0.f, e6 = ce ? ce[6] : 0.f, e7 = ce ? ce[7] : 0.f;
            s_gxy[cam][color][0] = -grx * (1.f - e4) + gry * e6;
            s_gxy[cam][color][1] = grx * e5 - gry * (1.f - e7);
        }
    }
    __syncthreads();
    if (threadIdx.x < 18) {
        const int q = threadIdx.x;
        const float v = s_red[0][q] + s_red[1][q] + s_red[2][q];
        o[1 + q] = v;   // C[6], kx[6], ky[6]
    } else if (threadIdx.x < 26) {
        const int c = (threadIdx.x - 18) >> 1, xyi = (threadIdx.x - 18) & 1;
        o[19 + (threadIdx.x - 18)] = s_gxy[c][0][xyi] + s_gxy[c][1][xyi] + s_gxy[c][2][xyi];
    } else if (threadIdx.x == 26) {
        float loss = 0.f;
        for (int p = 0; p < 6; p++) loss -= s_red[0][p] + s_red[1][p] + s_red[2][p];
        o[0] = loss;
    } else if (threadIdx.x < CORR_ADJ_STRIDE) {
        o[threadIdx.x] = 0.f;
    }
}

// Gradient of the correlation-peak loss by the extrinsics and the disparities, and the
// Gauss-Newton system of the LMA with the disparities eliminated (Schur complement).
// Workgroup = 32 tasks, 128 threads.  Per task t, with J_c the 2 x 14 derivatives of port c's
// centre (13 extrinsics, disparity), the loss gradient by the centres g_c and per pair p = (c1, c2)
// the curvatures kx, ky: g_t = sum_c g_c J_c, H_t = sum_p (J_c1 - J_c2)^T diag(kx, ky) (J_c1 - J_c2).
// With hd = H_t[13][13] (1 + lambda) + eps: A += H_ee - h h^T / hd, b += g_e - h g_d / hd,
// h = H_t[0..12][13].  Per workgroup partial sums of A (91 entries, upper triangle row-major),
// b (13), loss (1) and the plain extrinsic gradient sum_t g_e (13) go to part[block][118]; per task
// {g_d, hd, h[13]} to tstate[t][15].
constexpr int GEO_ADJ_TASKS = 32;
constexpr int GEO_ADJ_SUMS = 118;
__global__ __launch_bounds__(128) void k_geometry_adjoint(const PortGeomDev* __restrict__ pg, const double* __restrict__ rtab,
                                                          const int32_t* __restrict__ words, const int32_t* __restrict__ txy,
                                                          const float* __restrict__ disparity, const float* __restrict__ adj,
                                                          int num_tasks, double lambda, double* __restrict__ part,
                                                          double* __restrict__ tstate) {
    __shared__ double sJ[GEO_ADJ_TASKS][4][2][14];
    __shared__ double sG[GEO_ADJ_TASKS][4][14];
    __shared__ double sH13[GEO_ADJ_TASKS][14];
    __shared__ float sK[GEO_ADJ_TASKS][12];
    const int t0 = blockIdx.x * GEO_ADJ_TASKS;
    {   // derivatives of the 4 port centres of 32 tasks
        const int tl = threadIdx.x >> 2, c = threadIdx.x & 3, t = t0 + tl;
        double* rx = sJ[tl][c][0];
        double* ry = sJ[tl][c][1];
        const bool live = t < num_tasks && words[t] != 0 && !isnan(adj[(size_t)t * CORR_ADJ_STRIDE]);
        if (live) {
            const int xy16 = txy[t];
            const double px = (double)((xy16 & 0xffff) * 8 + 4), py = (double)(((xy16 >> 16) & 0xffff) * 8 + 4);
            double x, y;
            port_centre(*pg, rtab, px, py, (double)disparity[t], c, &x, &y, rx, ry);
            const float* a = adj + (size_t)t * CORR_ADJ_STRIDE;
            const double gx = a[19 + 2 * c], gy = a[20 + 2 * c];
            for (int k = 0; k < 14; k++) sG[tl][c][k] = gx * rx[k] + gy * ry[k];
            if (c == 0)
                for (int k = 0; k < 12; k++) sK[tl][k] = fmaxf(a[7 + k], 0.f);   // kx[6], ky[6]
        } else {
            for (int k = 0; k < 14; k++) rx[k] = ry[k] = sG[tl][c][k] = 0.0;
            if (c == 0)
                for (int k = 0; k < 12; k++) sK[tl][k] = 0.f;
        }
    }
    __syncthreads();
    auto hval = [&](int tl, int i, int j) {   // H_t[i][j]
        double s = 0.0;
#pragma unroll
        for (int p = 0; p < 6; p++) {
            const int c1 = QP_C1[p], c2 = QP_C2[p];
            const double dxi = sJ[tl][c1][0][i] - sJ[tl][c2][0][i], dxj = sJ[tl][c1][0][j] - sJ[tl][c2][0][j];
            const double dyi = sJ[tl][c1][1][i] - sJ[tl][c2][1][i], dyj = sJ[tl][c1][1][j] - sJ[tl][c2][1][j];
            s += (double)sK[tl][p] * dxi * dxj + (double)sK[tl][6 + p] * dyi * dyj;
        }
        return s;
    };
    for (int k = threadIdx.x; k < GEO_ADJ_TASKS * 14; k += 128) {
        const int tl = k / 14, i = k % 14;
        sH13[tl][i] = hval(tl, i, 13);
    }
    __syncthreads();
    const int e = threadIdx.x;
    if (e < GEO_ADJ_SUMS) {
        int ei = 0, ej = 0;
        if (e < 91) {   // upper triangle index -> (i, j)
            int r = e;
            while (r >= 13 - ei) { r -= 13 - ei; ei++; }
            ej = ei + r;
        }
        double acc = 0.0;
        for (int tl = 0; tl < GEO_ADJ_TASKS; tl++) {
            const int t = t0 + tl;
            if (t >= num_tasks) break;
            const double hd = sH13[tl][13] * (1.0 + lambda) + 1e-12;
            const double gd = sG[tl][0][13] + sG[tl][1][13] + sG[tl][2][13] + sG[tl][3][13];
            if (e < 91) {
                acc += hval(tl, ei, ej) - sH13[tl][ei] * sH13[tl][ej] / hd;
            } else if (e < 104) {
                const int i = e - 91;
                const double gi = sG[tl][0][i] + sG[tl][1][i] + sG[tl][2][i] + sG[tl][3][i];
                acc += gi - sH13[tl][i] * gd / hd;
            } else if (e == 104) {
                const float l = adj[(size_t)t * CORR_ADJ_STRIDE];
                if (!isnan(l)) acc += (double)l;
            } else {
                const int i = e - 105;
                acc += sG[tl][0][i] + sG[tl][1][i] + sG[tl][2][i] + sG[tl][3][i];
            }
        }
        part[(size_t)blockIdx.x * GEO_ADJ_SUMS + e] = acc;
    }
    for (int k = threadIdx.x; k < GEO_ADJ_TASKS; k += 128) {
        const int t = t0 + k;
        if (t >= num_tasks) break;
        double* ts = tstate + (size_t)t * 15;
        ts[0] = sG[k][0][13] + sG[k][1][13] + sG[k][2][13] + sG[k][3][13];
        ts[1] = sH13[k][13] * (1.0 + lambda) + 1e-12;
        for (int i = 0; i < 13; i++) ts[2 + i] = sH13[k][i];
    }
}

// Column sums of rows x cols doubles: one 256-thread workgroup per column (grid = cols), a strided
// partial sum per thread, then a tree in LDS (a fixed order: the same sums on every run).
__global__ __launch_bounds__(256) void k_sum_rows(const double* __restrict__ a, int rows, int cols, double* __restrict__ out) {
    __shared__ double s_part[256];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int r = threadIdx.x; r < rows; r += 256) s += a[(size_t)r * cols + c];
    s_part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_part[threadIdx.x] += s_part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = s_part[0];
}

// Correlation-peak offsets of the 6 quad pairs of each task (the measurement of the reference's
// extrinsic path: clt_mismatch, ID:2342-2420, by mismatchPairsCM, C2:1336-1430): per pair the
// integer maximum of its 15x15 correlation (the first in row order) and the centre of mass of the
// positive values within `radius` of it (image axes, +-7 px; the reference centres its window at the
// tile's residual disparity instead, the pairs here being aligned at the task disparity already).
// The peak sits at the excess of the pair's relative port offset (xy_c1 - xy_c2) over the aligned
// one, so the correction wanted is minus the peak.  Outputs (either may be null):
//  adj  a loss-adjoint record (CORR_ADJ_STRIDE floats) for k_geometry_adjoint, whose Gauss-Newton
//       step then solves min sum_p w_p |D_p delta + peak_p|^2 (D_p = J_c1 - J_c2): [0] sum w |peak|^2,
//       [7..12] and [13..18] the pair weights w_p (the correlation at the maximum less `floor`
//       where it exceeds the floor, else 0: the reference's tile weights, AC:1521-1525), [19..26]
//       per camera sum_{p: c1 = c} w_p peak_p - sum_{p: c2 = c} w_p peak_p;
//  offs [6][3] per task: dx, dy of the peak (NaN where no value in the window is positive) and the
//       correlation at the maximum (tasks without work: NaN, NaN, 0).
__global__ __launch_bounds__(256) void k_corr_offsets(const float* __restrict__ corr, const int32_t* __restrict__ tasks,
                                                      int task_words, int num_tasks, float radius, float floor,
                                                      float* __restrict__ adj, float* __restrict__ offs) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_tasks) return;
    const float qnan = __int_as_float(0x7fc00000);
    float* a = adj ? adj + (size_t)t * CORR_ADJ_STRIDE : nullptr;
    float* o = offs ? offs + (size_t)t * 18 : nullptr;
    if (tasks[(size_t)t * task_words] == 0) {
        if (a) {
            a[0] = qnan;
            for (int k = 1; k < CORR_ADJ_STRIDE; k++) a[k] = 0.f;
        }
        if (o)
            for (int p = 0; p < 6; p++) {
                o[3 * p] = qnan;
                o[3 * p + 1] = qnan;
                o[3 * p + 2] = 0.f;
            }
        return;
    }
    const int ir = (int)radius;
    const float r2 = radius * radius;
    float g[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float wp[6];
    float loss = 0.f;
    for (int p = 0; p < 6; p++) {
        const float* c = corr + ((size_t)t * 6 + p) * 225;
        int im = 0;
        float vm = c[0];
        for (int k = 1; k < 225; k++)
            if (c[k] > vm) { vm = c[k]; im = k; }
        const int iy = im / 15, ix = im % 15;
        float s0 = 0.f, sx = 0.f, sy = 0.f;
        for (int y = -ir; y <= ir; y++) {
            const int yy = iy + y;
            if (yy < 0 || yy > 14) continue;
            for (int x = -ir; x <= ir; x++) {
                const int xx = ix + x;
                if (xx < 0 || xx > 14 || (float)(x * x + y * y) > r2) continue;
                const float d = c[yy * 15 + xx];
                if (d > 0.f) {
                    s0 += d;
                    sx += d * (float)(xx - 7);
                    sy += d * (float)(yy - 7);
                }
            }
        }
        const bool ok = vm > floor && s0 > 0.f;
        const float w = ok ? vm - floor : 0.f;
        const float px = ok ? sx / s0 : 0.f, py = ok ? sy / s0 : 0.f;
        if (o) {
            o[3 * p] = s0 > 0.f ? sx / s0 : qnan;
            o[3 * p + 1] = s0 > 0.f ? sy / s0 : qnan;
            o[3 * p + 2] = vm;
        }
        wp[p] = w;
        loss += w * (px * px + py * py);
        g[2 * QP_C1[p]] += w * px;
        g[2 * QP_C1[p] + 1] += w * py;
        g[2 * QP_C2[p]] -= w * px;
        g[2 * QP_C2[p] + 1] -= w * py;
    }
    if (!a) return;
    a[0] = loss;
    for (int k = 0; k < 6; k++) a[1 + k] = 0.f;
    for (int k = 0; k < 6; k++) a[7 + k] = a[13 + k] = wp[k];
    for (int k = 0; k < 8; k++) a[19 + k] = g[k];
    for (int k = 27; k < CORR_ADJ_STRIDE; k++) a[k] = 0.f;
}

// Disparity step of the LMA after the extrinsic step de: dd_t = -(g_d + h . de) / hd, limited to
// max_step pixels (a tile whose correlations are flat along its disparity has a tiny hd).
__global__ __launch_bounds__(256) void k_disparity_step(const double* __restrict__ tstate, const double* __restrict__ de,
                                                        const float* __restrict__ d_in, float* __restrict__ d_out, int num_tasks,
                                                        double max_step) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_tasks) return;
    const double* ts = tstate + (size_t)t * 15;
    double s = ts[0];
    for (int i = 0; i < 13; i++) s += ts[2 + i] * de[i];
    const double dd = fmin(fmax(-s / ts[1], -max_step), max_step);
    d_out[t] = (float)((double)d_in[t] + dd);
}

// ====================================================================================
// textures: per tile RGBA (+ port weights, rms) from all cameras (ImageDtt.java:2662-2812,
// tile_debayer_shot_corr 3022-3147, tile_combine_rgba 2830-3000).  Block (256 threads) = tile,
// thread = pixel of the 16x16 tile in the pixel stages.  NC = number of cameras.
// ====================================================================================
// LDS layouts of a 16x16 texture tile (TexShared); both keep the 16-byte stores of a lane's four
// pixels (row n, columns 4g .. 4g+3) and the per-pixel reads of the combination free of bank
// conflicts.  LAY 0: rows 20 floats apart (the two lane groups of a half-wavefront on distinct
// banks).  LAY 1: rows 16 floats apart, the 16-byte column group g of row y stored at group
// g ^ ((y >> 1) & 3), 20 % less LDS per tile.
template <int LAY>
struct TexLayout {
    static constexpr int ROW = LAY ? 16 : 20, TILE = 16 * ROW;
    __device__ __forceinline__ static int at(int y, int x) {   // pixel (y, x)
        return LAY ? y * 16 + ((((x >> 2) ^ ((y >> 1) & 3)) << 2) | (x & 3)) : y * 20 + x;
    }
    __device__ __forceinline__ static int group(int y, int g) {   // pixels (y, 4g .. 4g+3)
        return LAY ? y * 16 + ((g ^ ((y >> 1) & 3)) << 2) : y * 20 + 4 * g;
    }
};

struct TexArgs {
    float col_w[3];
    float min_shot, scale_shot, ksigma, corr_magic_scale;
    float ksigma_l2;                  // ksigma * log2(e): weights as powers of 2
    float pair_k_l2[28];              // ksigma / |offset_j - offset_i|^2 * log2(e), pairs i < j
    int dust_remove, planes;
    float port_offsets[CLT_MAX_CAMS][2];
};


// Per-pixel port combination of one tile (tile_combine_rgba, ImageDtt.java:2830-3000, with the
// final mix ImageDtt.java:2789-2801).  ALL: every port of the NC is used (no per-port masking).
template <int NC, bool ALL, int TILE>
__device__ __forceinline__ void combine_pixel(const TexArgs& A, int port_mask, const float (&deb)[NC][3], float w2,
                                              float iw2, const float* s_nw, int pi, float* out, int i, int planes,
                                              float (&avg)[3]) {
    auto on = [&](int ip) { return ALL || ((port_mask >> ip) & 1); };
    int used = 0;
#pragma unroll
    for (int ip = 0; ip < NC; ip++) used += on(ip) ? 1 : 0;
    const float inv_used = fast_rcp((float)used);
    // Port weights are powers of 2 of scaled squared mismatches (exp(-k d) = 2^(-k log2(e) d)),
    // computed relative to the largest one (2^(m - a), m the smallest exponent): every ratio the
    // reference forms in double stays in float range; absolute weights are rescaled by 2^-m.
    float pw[NC], wscale = 1.f;
    avg[0] = avg[1] = avg[2] = 0.f;
#pragma unroll
    for (int ip = 0; ip < NC; ip++) pw[ip] = 0.f;
    if (used > 1) {
        constexpr int NP = NC * (NC - 1) / 2 > 0 ? NC * (NC - 1) / 2 : 1;
        float ea[NP];
        float m1 = 3.0e38f;
        int np = 0;
#pragma unroll
        for (int a = 0; a < NC; a++) {
#pragma unroll
            for (int b = a + 1; b < NC; b++) {
                float d = 0.f;
#pragma unroll
                for (int chn = 0; chn < 3; chn++) {
                    const float dc = deb[a][chn] - deb[b][chn];
                    d += A.col_w[chn] * dc * dc;
                }
                float e = A.pair_k_l2[np] * (d * iw2 * iw2);
                if (!ALL && !(on(a) && on(b))) e = 3.0e38f;
                m1 = fminf(m1, e);
                ea[np++] = e;
            }
        }
        np = 0;
#pragma unroll
        for (int a = 0; a < NC; a++) {
#pragma unroll
            for (int b = a + 1; b < NC; b++) {
                float e = fast_exp2(m1 - ea[np++]);
                if (!ALL && !(on(a) && on(b))) e = 0.f;
                pw[a] += e;
                pw[b] += e;
            }
        }
        // two best ports (written without dynamic register indexing)
        float bw1 = pw[0];
        int b1 = 0;
#pragma unroll
        for (int ip = 1; ip < NC; ip++)
            if (pw[ip] > bw1) { bw1 = pw[ip]; b1 = ip; }
        int b2 = (b1 == 0) ? 1 : 0;
        float bw2 = (b1 == 0) ? pw[1] : pw[0];
#pragma unroll
        for (int ip = 1; ip < NC; ip++)
            if (ip > b2 && ip != b1 && pw[ip] > bw2) { bw2 = pw[ip]; b2 = ip; }
        const float w1 = bw1 * fast_rcp(bw1 + bw2), wq = 1.f - w1;
#pragma unroll
        for (int chn = 0; chn < 3; chn++) {
            float v1 = 0.f, v2 = 0.f;
#pragma unroll
            for (int ip = 0; ip < NC; ip++) {
                v1 = (ip == b1) ? deb[ip][chn] : v1;
                v2 = (ip == b2) ? deb[ip][chn] : v2;
            }
            avg[chn] = w1 * v1 + wq * v2;
        }
        float a2[NC];
        float m2 = 3.0e38f;
        const float k2 = A.ksigma_l2 * iw2 * iw2;
#pragma unroll
        for (int ip = 0; ip < NC; ip++) {
            float d2 = 0.f;
#pragma unroll
            for (int chn = 0; chn < 3; chn++) {
                const float dc = deb[ip][chn] - avg[chn];
                d2 += A.col_w[chn] * dc * dc;
            }
            a2[ip] = k2 * d2;
            if (on(ip)) m2 = fminf(m2, a2[ip]);
        }
        float ksum = 0.f;
#pragma unroll
        for (int ip = 0; ip < NC; ip++) {
            pw[ip] = on(ip) ? fast_exp2(m2 - a2[ip]) : 0.f;
            ksum += pw[ip];
        }
        const float k = fast_rcp(ksum);
#pragma unroll
        for (int chn = 0; chn < 3; chn++) {
            float s = 0.f;
#pragma unroll
            for (int ip = 0; ip < NC; ip++) s += pw[ip] * deb[ip][chn];
            avg[chn] = k * s;
        }
        wscale = fast_exp2(-m2);
    } else if (used > 0) {
#pragma unroll
        for (int ip = 0; ip < NC; ip++)
            if (on(ip)) {
#pragma unroll
                for (int chn = 0; chn < 3; chn++) avg[chn] = deb[ip][chn];
                pw[ip] = 1.f;
            }
    }
    if (A.dust_remove && used == 4) {   // dust_remove (ImageDtt.java:3002-3019), scale free
        int wi = 0;
        float ww = pw[0];
#pragma unroll
        for (int ip = 1; ip < NC; ip++)
            if (pw[ip] < ww) { ww = pw[ip]; wi = ip; }
        float av = 0.f;
#pragma unroll
        for (int ip = 1; ip < NC; ip++) av += (ip != wi) ? pw[ip] : 0.f;
        av *= 1.f / (NC - 1);
        if (av > 0.f) {   // av == 0 only when those weights are below float range: they stay 0
            const float rav = fast_rcp(av);
            const float scale = 1.f + (av - ww) * rav * (1.f / (NC - 1));
            const float fw = ww * rav;
#pragma unroll
            for (int ip = 0; ip < NC; ip++) pw[ip] *= (ip == wi) ? fw : ((ip >= 1) ? scale : 1.f);
        }
    }
    float psum = 0.f;
#pragma unroll
    for (int ip = 0; ip < NC; ip++) psum += on(ip) ? pw[ip] : 0.f;
    out[3 * 256 + i] = psum * wscale * w2 * inv_used;
    float sw = 0.f;
#pragma unroll
    for (int ip = 0; ip < NC; ip++) sw += pw[ip];
    sw = (sw != 0.f) ? fast_rcp(sw) * w2 : 0.f;   // the window-free tiles times w2 are the IMCLT tiles
#pragma unroll
    for (int chn = 0; chn < 3; chn++) {
        float s = 0.f;
#pragma unroll
        for (int ip = 0; ip < NC; ip++) s += pw[ip] * s_nw[(ip * 3 + chn) * TILE + pi];
        out[chn * 256 + i] = sw * s;
    }
    if (planes > 4) {   // keep_weights: weights, then the rms of the ports (ImageDtt.java:2858-2874)
#pragma unroll
        for (int ip = 0; ip < NC; ip++) out[(4 + ip) * 256 + i] = pw[ip] * wscale;
        float swr = 0.f;
#pragma unroll
        for (int chn = 0; chn < 3; chn++) {   // population deviation, two-pass form (stable in float)
            float s1 = 0.f;
#pragma unroll
            for (int ip = 0; ip < NC; ip++)
                if (on(ip)) s1 += deb[ip][chn];
            const float mean = s1 * inv_used;
            float var = 0.f;
#pragma unroll
            for (int ip = 0; ip < NC; ip++)
                if (on(ip)) var += (deb[ip][chn] - mean) * (deb[ip][chn] - mean);
            const float r = fast_sqrt(var * inv_used);
            out[(4 + NC + chn) * 256 + i] = r;
            swr += A.col_w[chn] * r * r;
        }
        out[(4 + NC + 3) * 256 + i] = fast_sqrt(swr);
    }
}

// LDS of one texture tile: the window-free tiles (the IMCLT tiles divided by the squared window)
// and the debayered tiles, [unit][tile in layout LAY] (TexLayout, defined with TexArgs).
template <int NC, int LAY>
struct TexShared {
    static constexpr int NU = NC * 3;
    using Lay = TexLayout<LAY>;
    float nw[NU][Lay::TILE];
    float deb[NU][Lay::TILE];
    __device__ __forceinline__ float* deb_of(int u) { return &deb[u][0]; }
};

// Value of the lane one column to the left / right within the 16-lane row (DPP row shifts; 0 past
// the tile's first / last column, which is the zero border of the 3x3 debayer).  Every lane active.
__device__ __forceinline__ float unit_row_prev(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, false));
}
__device__ __forceinline__ float unit_row_next(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x101, 0xf, 0xf, false));
}
// value of lane (l + R) mod 16 of the lane's 16-lane row (DPP row_ror); the row's lanes active
template <int R>
__device__ __forceinline__ float row_ror(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x120 + R, 0xf, 0xf, false));
}
// value of lane (lane + d) mod 64 (LDS crossbar, no LDS memory); every lane active
__device__ __forceinline__ float lane_from(float v, int lane, int d) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(((lane + d) & 63) << 2, __float_as_int(v)));
}

// Inverse CLT of a unit on the matrix cores.  The 2-D inverse is out = U X U^T with X the 16x16
// coefficient matrix (rows (vs, kv), columns (hs, kh)) and U the 1-D inverse (Tables::iclt_u), two
// chains of four v_mfma_f32_16x16x4_f32 with lane (g, n) = (lane >> 4, lane & 15).  Pass 1
// (P = X^T U^T, contraction over (vs, kv)): A = X[(vs, kv)][(hs, kh) = n] with vs = g >> 1,
// kv = 4 (g & 1) + s for step s (four consecutive coefficients: one 16-byte load), B = U[y = n][4g + s];
// P lands as rows (hs, kh) = 4g + r, column y = n.  Pass 2 (out^T = U P, contraction over (hs, kh))
// takes P's register s as its B operand for step s (the contraction order follows the rows each
// lane holds, so no data moves between the passes), with A = U[x = n][4g + s], the same constants;
// pixel (y = n, x = 4g + r) lands in register r.
// Per-lane constants of the texture units (set once per block): the matrix operand, LPF * 0.25 of
// the lane's coefficients (kh = n & 7, kv = 4 (g & 1) + s), and for its pixels the inverse squared
// window, the squared window, and the renormalisation of the 3x3 debayer kernels for the taps
// dropped at the tile border (ImageDtt.java:43-50, ID:3048-3147: red/blue 4/3 per border
// dimension, green 8/7 at a side and 4/3 at a corner).
struct TexLane {
    float u[4], lpf[4], nwf[4], w2f[4];
    bool redge, cedge_lo, cedge_hi;   // the lane's row / its register 0 / register 3 on the tile border
};
__device__ __forceinline__ void tex_lane_setup(TexLane& L, const float* __restrict__ lpf) {
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    const int qy = n < 8 ? n : 15 - n;
    const bool redge = n == 0 || n == 15;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        L.u[s] = c_tab.iclt_u[n][4 * g + s];
        L.lpf[s] = 0.25f * lpf[(n & 7) * 8 + 4 * (g & 1) + s];
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int x = 4 * g + r, qx = x < 8 ? x : 15 - x;
        L.nwf[r] = c_tab.ihw2[qy] * c_tab.ihw2[qx];
        L.w2f[r] = c_tab.hw2[qy] * c_tab.hw2[qx];
    }
    L.redge = redge;
    L.cedge_lo = g == 0;
    L.cedge_hi = g == 3;
}
// renormalisation of the debayer kernel at pixel (n, 4g + r) of lane L (see TexLane)
__device__ __forceinline__ float deb_factor(const TexLane& L, int r, bool green) {
    const bool cedge = (r == 0 && L.cedge_lo) || (r == 3 && L.cedge_hi);
    if (green) return L.redge ? (cedge ? 4.f / 3.f : 8.f / 7.f) : (cedge ? 8.f / 7.f : 1.f);
    return (L.redge ? 4.f / 3.f : 1.f) * (cedge ? 4.f / 3.f : 1.f);
}

// IMCLT of the NC*3 camera/colour units of one tile after LPF * 0.25 (ImageDtt.java:2684-2716),
// with the optional fract_shift by the extra disparity (ID:2663-2677), then the debayer with
// shot-noise correction and the re-applied window (tile_debayer_shot_corr, ID:3022-3147).  Wavefront
// w takes units w, w + 4, ... (all their loads first, then their matrix-core chains interleaved);
// the unit's colour is uniform in the wavefront.  Lane (g, n) holds pixels (n, 4g .. 4g+3): the
// vertical debayer taps come from the neighbouring lanes by DPP, the horizontal ones from the
// lane's own registers and, at its first / last column, from the lane group to the left / right.
// x: the coefficients from tex_fetch; next() is called once they are consumed.  Leaves S.nw and the debayered tiles in LDS; sink(uu, j, px)
// receives the windowed IMCLT tile of unit uu (the wavefront's j-th unit) while it is in registers.
// All 256 threads call this.
// Coefficients of the wavefront's units of one tile (see tex_units): lane (g, n) reads mode
// (n >> 3) | ((g >> 1) << 1), kh = n & 7, kv = 4 (g & 1) .. + 3 (one 16-byte load per unit).
template <int NC>
__device__ __forceinline__ void tex_fetch(const float* __restrict__ clt, size_t tile, size_t ntiles,
                                          float (&x)[(NC * 3 + 3) / 4][4]) {
    constexpr int NU = NC * 3, J = (NU + 3) / 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
#pragma unroll
    for (int j = 0; j < J; j++) {
        const int uu = min(wave + 4 * j, NU - 1);   // a wavefront without a j-th unit repeats its last
        const int cam = uu / 3, color = uu - 3 * cam;
        const float4 v = *reinterpret_cast<const float4*>(clt + (((size_t)cam * ntiles + tile) * 3 + color) * 256 +
                                                         (n >> 3) * 64 + (g >> 1) * 128 + (n & 7) * 8 + 4 * (g & 1));
        x[j][0] = v.x; x[j][1] = v.y; x[j][2] = v.z; x[j][3] = v.w;
    }
}

template <int NC, int LAY, class Sink, class Next>
__device__ __forceinline__ void tex_units(float (&x)[(NC * 3 + 3) / 4][4], const TexLane& L, const TexArgs& A,
                                          float extra, bool do_shift, TexShared<NC, LAY>& S, Sink sink, Next next) {
    constexpr int NU = NC * 3, J = (NU + 3) / 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    const int hs = n >> 3, kh = n & 7;
    if (do_shift) {   // fract_shift by extra * port_offset / corr_magic_scale (ID:2663-2677)
        const bool vodd = g >= 2;   // vs = 1
        const float sigma = hs ? 1.f : -1.f;
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int cam = min(wave + 4 * j, NU - 1) / 3;
            const float sxp = extra * A.port_offsets[cam][0] / A.corr_magic_scale;
            const float syp = extra * A.port_offsets[cam][1] / A.corr_magic_scale;
            float sh, ch;
            sincos_pi((kh + 0.5f) * sxp * 0.125f, &sh, &ch);
#pragma unroll
            for (int s = 0; s < 4; s++) {
                // horizontal: the other horizontal mode is lane n ^ 8; vertical: lane (lane + 32) mod 64
                const float t = x[j][s] * ch + lane_xor8(x[j][s]) * sigma * sh;
                const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(t), __float_as_uint(t), false, false);
                const float tw = __uint_as_float(vodd ? sw[0] : sw[1]);
                float sv, cv;
                sincos_pi((4 * (g & 1) + s + 0.5f) * syp * 0.125f, &sv, &cv);
                x[j][s] = vodd ? t * cv + tw * sv : t * cv - tw * sv;
            }
        }
    }
    float4x d1[J], d2[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        d1[j] = (float4x){0.f, 0.f, 0.f, 0.f};
        d2[j] = (float4x){0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
        for (int j = 0; j < J; j++) d1[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[j][s] * L.lpf[s], L.u[s], d1[j], 0, 0, 0);
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
        for (int j = 0; j < J; j++) d2[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(L.u[s], d1[j][s], d2[j], 0, 0, 0);
    next();   // x is free again: the caller may load the next tile's coefficients into it
    const float kshot = fast_rcp(fast_sqrt(A.min_shot));
#pragma unroll
    for (int j = 0; j < J; j++) {
        const int uu = wave + 4 * j;
        if (uu >= NU) break;   // wave-uniform
        const int color = uu % 3;
        float px[4] = {d2[j][0], d2[j][1], d2[j][2], d2[j][3]};
        sink(uu, j, px);
        float nw[4];
#pragma unroll
        for (int r = 0; r < 4; r++) nw[r] = px[r] * L.nwf[r];
        *reinterpret_cast<float4*>(&S.nw[uu][TexLayout<LAY>::group(n, g)]) = make_float4(nw[0], nw[1], nw[2], nw[3]);
        // horizontal neighbour sums (columns 4g - 1 and 4g + 4 from the lane groups to the left / right)
        const float left = lane_from(nw[3], lane, 48), right = lane_from(nw[0], lane, 16);
        float hsum[4];
        hsum[0] = (g > 0 ? left : 0.f) + nw[1];
        hsum[1] = nw[0] + nw[2];
        hsum[2] = nw[1] + nw[3];
        hsum[3] = nw[2] + (g < 3 ? right : 0.f);
        const bool green = color == 2;   // wave-uniform
        float dv[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float vsum = unit_row_prev(nw[r]) + unit_row_next(nw[r]);
            const float side = vsum + hsum[r];
            if (green) {   // kern_g: centre 1/2, sides 1/8
                dv[r] = deb_factor(L, r, true) * (0.5f * nw[r] + 0.125f * side);
            } else {       // kern_rb: centre 1/4, sides 1/8, corners 1/16
                const float corner = unit_row_prev(hsum[r]) + unit_row_next(hsum[r]);
                dv[r] = deb_factor(L, r, false) * (0.25f * nw[r] + 0.125f * side + 0.0625f * corner);
            }
            if (A.scale_shot > 0.f) dv[r] = A.scale_shot * ((dv[r] > A.min_shot) ? fast_sqrt(dv[r]) : (kshot * dv[r]));
            dv[r] *= L.w2f[r];
        }
        *reinterpret_cast<float4*>(&S.deb[uu][TexLayout<LAY>::group(n, g)]) = make_float4(dv[0], dv[1], dv[2], dv[3]);
    }
}

// Port combination (pixel per thread) and the tile statistics of one texture tile, from the LDS
// that tex_units left (the caller's barrier between them).  All 256 threads call this.
template <int NC, int LAY>
__device__ __forceinline__ void tex_combine(const TexArgs& A, int port_mask, TexShared<NC, LAY>& S, float* __restrict__ out,
                                            float* __restrict__ stats_t) {
    const int i = threadIdx.x;   // pixel index in the 16x16 tile
    const int planes = A.planes;
    const int py = i >> 4, px = i & 15;
    const int pi = TexLayout<LAY>::at(py, px);   // ... in the LDS tiles
    const int qy = py < 8 ? py : 15 - py, qx = px < 8 ? px : 15 - px;
    const float w2 = c_tab.hw2[qy] * c_tab.hw2[qx];
    const float iw2 = c_tab.ihw2[qy] * c_tab.ihw2[qx];
    float deb[NC][3];
#pragma unroll
    for (int uu = 0; uu < NC * 3; uu++) deb[uu / 3][uu % 3] = S.deb_of(uu)[pi];
    float avg[3];
    constexpr int TILE = TexLayout<LAY>::TILE;
    if (port_mask == (1 << NC) - 1) combine_pixel<NC, true, TILE>(A, port_mask, deb, w2, iw2, &S.nw[0][0], pi, out, i, planes, avg);
    else combine_pixel<NC, false, TILE>(A, port_mask, deb, w2, iw2, &S.nw[0][0], pi, out, i, planes, avg);
    if (stats_t) {   // max_diff (ImageDtt.java:2965-2981) and ports_rgb (ImageDtt.java:2982-2996)
        // transpose through LDS: value k of pixel i at [k][i] (the NC max_diff rows reuse deb, the
        // 3*NC colour rows reuse nw); 16 threads then reduce each row
        float* red_max = &S.deb[0][0];
        float* red_rgb = &S.nw[0][0];
        // rows RS floats apart: 272 puts the rows read by the two 16-lane halves of a half-wavefront
        // on distinct banks (256 would put all four rows of a wavefront on the same banks); it fits
        // the tiles of the padded layout
        constexpr int RS = (TexLayout<LAY>::TILE >= 272) ? 272 : 256;
        __syncthreads();
#pragma unroll
        for (int ip = 0; ip < NC; ip++) {
            float d2 = 0.f;
            if ((port_mask >> ip) & 1) {
#pragma unroll
                for (int chn = 0; chn < 3; chn++) {
                    const float dc = deb[ip][chn] - avg[chn];
                    d2 += dc * dc * A.col_w[chn];
                }
                d2 *= w2;
            }
            red_max[ip * RS + i] = d2;
#pragma unroll
            for (int chn = 0; chn < 3; chn++) red_rgb[(chn * NC + ip) * RS + i] = deb[ip][chn];
        }
        __syncthreads();
        // 4*NC values x 256 pixels: 16 threads per value, each the 16 pixels j, j + 16, ... (the
        // 16 threads read consecutive words), then a reduction over the 16-lane row by DPP row
        // rotations; a value's kind (maximum or mean) is uniform in its 16 lanes, and in the whole
        // wavefront when NC is a multiple of 4
        for (int kb = 0; kb < 4 * NC; kb += 16) {
            const int k = kb + (i >> 4), j = i & 15;
            if (k < 4 * NC) {
                const bool is_max = k < NC;
                const float* r = (is_max ? red_max + k * RS : red_rgb + (k - NC) * RS) + j;
                float v = r[0];
                if (is_max) {
#pragma unroll
                    for (int q = 1; q < 16; q++) v = fmaxf(v, r[16 * q]);
                    v = fmaxf(v, row_ror<8>(v));
                    v = fmaxf(v, row_ror<4>(v));
                    v = fmaxf(v, row_ror<2>(v));
                    v = fmaxf(v, row_ror<1>(v));
                    if (j == 0) stats_t[k] = sqrtf(v);
                } else {
#pragma unroll
                    for (int q = 1; q < 16; q++) v += r[16 * q];
                    v += row_ror<8>(v);
                    v += row_ror<4>(v);
                    v += row_ror<2>(v);
                    v += row_ror<1>(v);
                    if (j == 0) stats_t[k] = v * (1.f / 256.f);
                }
            }
        }
    }
}

// NaN texture tile and statistics of a task without work (flags 0)
template <int NC>
__device__ __forceinline__ void tex_nan(float* __restrict__ out, float* __restrict__ stats_t, int planes) {
    const float qnan = __int_as_float(0x7fc00000);
    for (int p = 0; p < planes; p++) out[p * 256 + threadIdx.x] = qnan;
    if (stats_t && threadIdx.x < 4 * NC) stats_t[threadIdx.x] = qnan;
}

struct NoRowSink {
    __device__ __forceinline__ void operator()(int, int, const float*) const {}
};

template <int NC>
__global__ __launch_bounds__(256) void k_textures(const float* __restrict__ clt, const int32_t* __restrict__ tasks,
                                                  int task_words, int num_tasks, const float* __restrict__ lpf,
                                                  const float* __restrict__ extra_disp, const TexArgs* __restrict__ args_p,
                                                  float* __restrict__ tiles_out, float* __restrict__ stats_out,
                                                  const int32_t* __restrict__ task_frame, long long clt_fstride, GeomDev g) {
    __shared__ __attribute__((aligned(16))) TexShared<NC, 1> S;
    const TexArgs& A = *args_p;
    const int t = xcd_block(blockIdx.x, gridDim.x);
    if (t >= num_tasks) return;
    const int32_t* rec = tasks + (size_t)t * task_words;
    const int flags = rec[0];
    float* out = tiles_out + (size_t)t * A.planes * 256;
    float* stats_t = stats_out ? stats_out + (size_t)t * 4 * NC : nullptr;
    if (flags == 0) {
        tex_nan<NC>(out, stats_t, A.planes);
        return;
    }
    const int port_mask = flags & 0xff & ((1 << NC) - 1);
    const int txy = rec[1];
    const size_t tile = (size_t)((txy >> 16) & 0xffff) * g.tiles_x + (txy & 0xffff);
    const size_t ntiles = (size_t)g.tiles_x * g.tiles_y;
    if (task_frame) clt += task_frame[t] * clt_fstride;   // batched frames
    float extra = extra_disp ? extra_disp[t] : 0.f;
    if (isnan(extra)) extra = 0.f;
    const bool do_shift = (extra != 0.f) && !((flags >> 24) & 1);
    float x[(NC * 3 + 3) / 4][4];
    tex_fetch<NC>(clt, tile, ntiles, x);
    TexLane L;
    tex_lane_setup(L, lpf);
    tex_units<NC>(x, L, A, extra, do_shift, S, NoRowSink{}, [] {});
    __syncthreads();
    tex_combine<NC, 1>(A, port_mask, S, out, stats_t);
}

// ====================================================================================
// textures + imclt_rbg in one pass over the converted tiles: each tile's NC*3 units go through
// the IMCLT once; the windowed IMCLT tiles feed both the texture tile (tex_units / tex_combine) and
// the overlap-add of the (W+8) x (H+8) image planes (imclt_rbg, same LPF * 0.25).
// A block sweeps one tile column down a segment of seg_rows output block rows, starting one tile
// above the segment (its bottom half only); the lower half of each unit's tile (lanes n >= 8) waits
// in LDS until the next tile adds its upper half (lanes n < 8).  The two halves
// of a tile's 16 columns overlap the neighbouring columns, so the columns run in two launches:
// PAR = 0 (even tx) stores its 16 output columns, PAR = 1 (odd tx) adds to them; when tiles_x is
// even the last column's right half has no even neighbour and is stored.  The even launch reads no
// image planes, so with the 16-float LDS rows (LAY 1) it fits 96 VGPRs and runs 5 waves per SIMD.
// Requires NC * 3 <= 16, no extra disparity, and at most one task per tile with every task's flags
// non-zero (tile_map then names every task exactly once).
// ====================================================================================
template <int NC, int LAY, int PAR>
__global__ __launch_bounds__(256, NC <= 4 ? ((PAR == 0 && LAY == 1) ? 5 : 4) : 3) void k_tex_imclt(const float* __restrict__ clt, const int32_t* __restrict__ tasks,
                                                   int task_words, const int32_t* __restrict__ tile_map,
                                                   const float* __restrict__ lpf, const TexArgs* __restrict__ args_p,
                                                   float* __restrict__ tiles_out, float* __restrict__ stats_out,
                                                   float* __restrict__ rbg, int seg_rows, GeomDev g,
                                                   long long clt_fstride) {
    static_assert(NC * 3 <= 16, "at most four units per wavefront");
    constexpr int NU = NC * 3, J = (NU + 3) / 4;
    __shared__ __attribute__((aligned(16))) TexShared<NC, LAY> S;
    const TexArgs& A = *args_p;
    const int tx = 2 * (int)blockIdx.x + PAR;
    if (tx >= g.tiles_x) return;   // whole block
    const int by0 = blockIdx.y * seg_rows;
    const int nby = min(seg_rows, g.tiles_y + 1 - by0);
    const size_t ntiles = (size_t)g.tiles_x * g.tiles_y;
    const int out_w = g.width + 8, out_h = g.height + 8;
    const size_t plane_sz = (size_t)out_w * out_h;
    // frame blockIdx.z of a batch: its CLT, tile map (of flat task indices) and image planes
    clt += blockIdx.z * clt_fstride;
    tile_map += blockIdx.z * ntiles;
    rbg += blockIdx.z * (NC * 3) * plane_sz;
    TexLane L;
    tex_lane_setup(L, lpf);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63, gq = lane >> 4, n = lane & 15;
    // right half stored, not added: the last column when no even column follows it
    const bool add_left = PAR == 1;
    const bool add_right = PAR == 1 && !(tx == g.tiles_x - 1 && (g.tiles_x & 1) == 0);
    const bool add_mine = (gq < 2) ? add_left : add_right;   // columns 4 gq .. 4 gq + 3 of the tile
    // lower halves (rows 8-15) of the previous tile's units, [unit][row - 8][lane group]: written by
    // lanes n >= 8, read by lanes n < 8 of the same wavefront one tile later (the first tile of a
    // segment is never stored, so the array needs no clearing)
    __shared__ float4 s_carry[NU][8][4];
    // coefficients of the next tile down the column, loaded one tile ahead (rows outside the frame
    // read a clamped row and are not used)
    auto row_tile = [&](int ty) { return (size_t)min(max(ty, 0), g.tiles_y - 1) * g.tiles_x + tx; };
    float x[J][4];
    tex_fetch<NC>(clt, row_tile(by0 - 1), ntiles, x);
    int tn = (by0 >= 1) ? tile_map[row_tile(by0 - 1)] : -1;
    for (int ty = by0 - 1; ty < by0 + nby; ty++) {
        const int t = tn;
        const bool more = ty + 1 < by0 + nby;
        tn = (more && ty + 1 < g.tiles_y) ? tile_map[row_tile(ty + 1)] : -1;
        auto next = [&] {
            if (more) tex_fetch<NC>(clt, row_tile(ty + 1), ntiles, x);
        };
        const bool own = ty >= by0;   // the segment writes output block row ty
        // pixels (n, 4 gq .. 4 gq + 3) of the upper half (n < 8): output row ty * 8 + n, columns tx * 8 + 4 gq ..
        const size_t o0 = (size_t)(max(ty, 0) * 8 + (n & 7)) * out_w + tx * 8 + 4 * gq;
        // the odd columns' read of the even columns' sums, issued before this tile's transforms
        float4 pre[J];
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int uu = min(wave + 4 * j, NU - 1);
            pre[j] = (own && n < 8 && add_mine && wave + 4 * j < NU)
                         ? *reinterpret_cast<const float4*>(rbg + uu * plane_sz + o0)
                         : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        auto sink = [&](int uu, int j, const float* px) {
            // the lower half of the tile above, read before this tile's lower half replaces it
            const float4 up = s_carry[uu][n & 7][gq];
            if (own && n < 8)
                *reinterpret_cast<float4*>(rbg + uu * plane_sz + o0) =
                    make_float4(px[0] + up.x + pre[j].x, px[1] + up.y + pre[j].y, px[2] + up.z + pre[j].z,
                                px[3] + up.w + pre[j].w);
            if (n >= 8) s_carry[uu][n & 7][gq] = make_float4(px[0], px[1], px[2], px[3]);
        };
        if (t >= 0) {
            const int32_t* rec = tasks + (size_t)t * task_words;
            const int port_mask = rec[0] & 0xff & ((1 << NC) - 1);
            tex_units<NC>(x, L, A, 0.f, false, S, sink, next);
            __syncthreads();
            if (own && ty < g.tiles_y)
                tex_combine<NC, LAY>(A, port_mask, S, tiles_out + (size_t)t * A.planes * 256,
                                stats_out ? stats_out + (size_t)t * 4 * NC : nullptr);
            __syncthreads();   // the next tile reuses the LDS
        } else {   // no task tile (or outside the frame): nothing to add but the lower half above
            next();
            const float zero[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int j = 0; j < J; j++)
                if (wave + 4 * j < NU) sink(wave + 4 * j, j, zero);
        }
    }
}

// ====================================================================================
// combine textures: overlap-add of the texture tiles (ImageDtt.java:3858-3989, overlap,
// sharp_alpha = false), gather form: thread = output pixel.
// ====================================================================================
// combineRGBATiles (ImageDtt.java:3855-3991) as a per-pixel gather over the tile map.  overlap: the
// (up to 4) 16x16 tiles at stride 8 covering the pixel are summed; with sharp_alpha the alpha plane
// comes only from the centre 8x8 of the tile containing the pixel.  Without overlap the output is the
// (2 height) x (2 width) mosaic of whole tiles.  Every output pixel is written once.
__global__ __launch_bounds__(256) void k_combine_textures(const float* __restrict__ tiles, const int32_t* __restrict__ tile_map,
                                                          int planes, int overlap, int sharp_alpha,
                                                          float* __restrict__ out, GeomDev g) {
    const int ow = overlap ? g.width : 2 * g.width, oh = overlap ? g.height : 2 * g.height;
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= ow || y >= oh) return;
    const size_t plane_sz = (size_t)ow * oh, o = (size_t)y * ow + x;
    if (!overlap) {
        const int tk = tile_map[(size_t)(y >> 4) * g.tiles_x + (x >> 4)];
        const int off = (y & 15) * 16 + (x & 15);
        for (int p = 0; p < planes; p++) out[p * plane_sz + o] = (tk < 0) ? 0.f : tiles[((size_t)tk * planes + p) * 256 + off];
        return;
    }
    const int tyh = (y + 4) >> 3, txh = (x + 4) >> 3;
    const int iy = (y + 4) & 7, ix = (x + 4) & 7;
    int tsk[4], off[4];
    int n = 0;
#pragma unroll
    for (int a = 0; a < 2; a++) {
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int ty = tyh - a, tx = txh - b;
            if (ty < 0 || ty >= g.tiles_y || tx < 0 || tx >= g.tiles_x) continue;
            const int tk = tile_map[(size_t)ty * g.tiles_x + tx];
            if (tk < 0) continue;
            tsk[n] = tk;
            off[n] = (iy + 8 * a) * 16 + ix + 8 * b;
            n++;
        }
    }
    for (int p = 0; p < planes; p++) {
        float s = 0.f;
        if (p == 3 && sharp_alpha) {   // the centre 8x8 of the pixel's own tile
            const int tk = tile_map[(size_t)(y >> 3) * g.tiles_x + (x >> 3)];
            if (tk >= 0) s = tiles[((size_t)tk * planes + 3) * 256 + ((y & 7) + 4) * 16 + (x & 7) + 4];
        } else {
            for (int k = 0; k < n; k++) s += tiles[((size_t)tsk[k] * planes + p) * 256 + off[k]];
        }
        out[p * plane_sz + o] = s;
    }
}

}  // namespace clt
