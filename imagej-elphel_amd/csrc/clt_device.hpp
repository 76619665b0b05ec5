// clt_device.hpp -- device-side building blocks of the CLT tile processor (gfx950).
//
// Work unit of the transform kernels: one 8x8 tile of one camera and one colour, handled by a
// group of 16 lanes (4 units per 64-lane wavefront).  Separable 8-point transforms run in
// registers (one 8-vector per lane), the transposes between the horizontal and the vertical pass
// go through a 1 KiB LDS scratch per unit, partner exchanges between the two horizontal modes use
// cross-lane shuffles.  CLT coefficients are stored [mode][kh][kv] (horizontal frequency major,
// the order of the reference CPU path, DttRad2.java:503-519).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clt {

// Tables computed on the host in double precision from their closed forms and uploaded once.
struct Tables {
    float hw[8];        // half window sin(pi*(i+0.5)/16)            (DttRad2.java:684)
    float hw2[8];       // hw^2 (one factor of the squared 2-D window, ImageDtt.java:2686-2690)
    float ihw2[8];      // 1 / hw^2
    float d4pre[4][2];  // 0.5 * exp(-i pi (n + 1/4) / 8): DCT-IV pre-twiddle (with the 1/sqrt(4) scale)
    float d4post[4][2]; // exp(-i pi k / 8): DCT-IV post-twiddle
    float ciie[8][8];   // DCT-IIe                                    (DttRad2.java:952-968)
    float siie[8][8];   // DST-IIe                                    (DttRad2.java:1014-1030)
    // 1-D inverse lapped transform as a 16x16 matrix: sample t of the 16-long windowed output from
    // mode m (0: DCT-IV, 1: DST-IV) coefficient k at column m*8+k (the IDCT-IV / IDST-IV, unfold and
    // window of DttRad2.java:392-434), the operand of the matrix-core inverse CLT
    float iclt_u[16][16];
};
extern __constant__ Tables c_tab;

typedef float float4x __attribute__((ext_vector_type(4)));   // v_mfma_f32_16x16x4_f32 accumulator

struct GeomDev {
    int width, height, cams, colors, red_row, red_col;
    int kh, kv, kstep;            // kernel grid and step
    int tiles_x, tiles_y;
};

// ---------------------------------------------------------------- lane / wave helpers
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Value of lane (i ^ 8) within each row of 16 lanes: DPP row rotate by 8 (a VALU move, no LDS).
// Every lane of the wavefront must be active.
__device__ __forceinline__ float lane_xor8(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xf, 0xf, false));
}

// Hardware transcendentals (about 1 ulp, no denormal range handling): used where the argument
// range is known to be normal and the reference's double result is matched to ~1e-6 anyway.
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_log2(float x) { return __builtin_amdgcn_logf(x); }

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of the
// logical block space so neighbouring tiles share an L2.
__device__ __forceinline__ int xcd_block(int bid, int nblocks) {
    const int nx = 8;
    int per = nblocks / nx, rem = nblocks % nx;
    int x = bid % nx, local = bid / nx;
    return (x < rem) ? x * (per + 1) + local : rem * (per + 1) + (x - rem) * per + local;
}

// Order of a task list over the blocks: the list is cut into bands of B = TASK_BAND rows of a dense
// row-major list W tiles wide (positions pBW .. pBW + BW - 1), and a band is taken column by
// column, so the tiles that share an aberration kernel tile (kernel grid at twice the tile step,
// above and beside each other) and window rows run within a few blocks of each other, while those
// are in L2; in row order the tile row below comes a whole row later, after they have left it.  A
// bijection on 0 .. n-1 (positions past the last whole band keep their place), so any list is
// still covered exactly once.
constexpr int TASK_BAND = 8;
__device__ __forceinline__ int band_order(int i, int n, int w) {
    const unsigned gsz = TASK_BAND * (unsigned)w;
    const unsigned p = (unsigned)i / gsz;
    if ((p + 1) * gsz > (unsigned)n) return i;
    const unsigned q = (unsigned)i - p * gsz;
    return (int)(p * gsz + (q % TASK_BAND) * (unsigned)w + q / TASK_BAND);
}

// ---------------------------------------------------------------- 8-point transforms
// 8-point orthonormal DCT-IV (its own inverse) through a 4-point complex DFT between two twiddle
// stages: t_n = (x[2n] + i x[7-2n]) pre_n, T = DFT4(t), u_k = T_k post_k, y[2k] = Re u_k,
// y[7-2k] = -Im u_k (45 operations instead of the 64 of the matrix product).  REV: the input is
// read reversed and the odd outputs negated, which is DST-IV(x) = (-1)^k DCT-IV(reverse x)[k]
// (DttRad2.java:88-98).
template <bool REV>
__device__ __forceinline__ void dtt_iv8(const float* x, float* y) {
    float tr[4], ti[4];
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const float a = REV ? x[7 - 2 * n] : x[2 * n];
        const float b = REV ? x[2 * n] : x[7 - 2 * n];
        const float c = c_tab.d4pre[n][0], s = c_tab.d4pre[n][1];
        tr[n] = a * c - b * s;
        ti[n] = a * s + b * c;
    }
    const float ar = tr[0] + tr[2], ai = ti[0] + ti[2], br = tr[0] - tr[2], bi = ti[0] - ti[2];
    const float cr = tr[1] + tr[3], ci = ti[1] + ti[3], dr = tr[1] - tr[3], di = ti[1] - ti[3];
    const float Tr[4] = {ar + cr, br + di, ar - cr, br - di};
    const float Ti[4] = {ai + ci, bi - dr, ai - ci, bi + dr};
    y[0] = Tr[0];
    y[7] = REV ? Ti[0] : -Ti[0];
#pragma unroll
    for (int k = 1; k < 4; k++) {
        const float wr = c_tab.d4post[k][0], wi = c_tab.d4post[k][1];
        const float ur = Tr[k] * wr - Ti[k] * wi;
        const float ui = Tr[k] * wi + Ti[k] * wr;
        y[2 * k] = ur;
        y[7 - 2 * k] = REV ? ui : -ui;
    }
}
__device__ __forceinline__ void dct_iv8(const float* x, float* y) { dtt_iv8<false>(x, y); }
__device__ __forceinline__ void dst_iv8(const float* x, float* y) { dtt_iv8<true>(x, y); }

// ---------------------------------------------------------------- fold / unfold index algebra
// DttRad2.java:112-144 get_fold_indices(x, 8): tap b of folded output x reads source index
// fold_src, window hw[fold_w] (+ derivative term hw[fold_dw] * fold_ds for a shifted window),
// with sign fold_sc for the cosine (C) and fold_ss for the sine (S) fold.
__host__ __device__ constexpr int fold_src(int x, int b) {
    return (x < 4) ? (b == 0 ? 11 - x : 12 + x) : (b == 0 ? x - 4 : 11 - x);
}
__host__ __device__ constexpr int fold_w(int x, int b) {
    return (x < 4) ? (b == 0 ? 4 + x : 3 - x) : (b == 0 ? x - 4 : 11 - x);
}
__host__ __device__ constexpr int fold_dw(int x, int b) {
    return (x < 4) ? (b == 0 ? 3 - x : 4 + x) : (b == 0 ? 11 - x : x - 4);
}
__host__ __device__ constexpr int fold_ds(int x, int b) { return (x < 4) ? -1 : 1; }
__host__ __device__ constexpr int fold_sc(int x, int b) { return (x < 4) ? -1 : (b == 0 ? 1 : -1); }
__host__ __device__ constexpr int fold_ss(int x, int b) { return (x < 4) ? (b == 0 ? 1 : -1) : 1; }

// DttRad2.java:392-403 get_unfold_index_signs(x, 8): output x of the 16-long inverse reads IDCT
// index unfold_idx, signs unfold_sc (C) / unfold_ss (S), window hw[x < 8 ? x : 15 - x].
__host__ __device__ constexpr int unfold_idx(int x) {
    return (x < 4) ? x + 4 : (x < 8) ? 11 - x : (x < 12) ? 11 - x : x - 12;
}
__host__ __device__ constexpr int unfold_sc(int x) { return (x < 4) ? 1 : -1; }
__host__ __device__ constexpr int unfold_ss(int x) { return (x < 12) ? 1 : -1; }
__host__ __device__ constexpr int unfold_w(int x) { return (x < 8) ? x : 15 - x; }

// Shifted-window lapped fold of a 16-vector into the cosine and sine 8-vectors
// (DttRad2.java:263-314 get_shifted_fold_2d, one dimension): ac, as = cos, sin(pi/16 * shift).
__device__ __forceinline__ void fold16(const float* x, float ac, float as, float* yc, float* ys) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
        float w0 = fmaf(as, c_tab.hw[fold_dw(i, 0)] * fold_ds(i, 0), ac * c_tab.hw[fold_w(i, 0)]);
        float w1 = fmaf(as, c_tab.hw[fold_dw(i, 1)] * fold_ds(i, 1), ac * c_tab.hw[fold_w(i, 1)]);
        float a = w0 * x[fold_src(i, 0)], b = w1 * x[fold_src(i, 1)];
        yc[i] = fold_sc(i, 0) * a + fold_sc(i, 1) * b;
        ys[i] = fold_ss(i, 0) * a + fold_ss(i, 1) * b;
    }
}

// Inverse lapped unfold of the cosine/sine IDCT outputs to 16 samples (DttRad2.java:405-434).
__device__ __forceinline__ void unfold16(const float* oc, const float* os, float* y) {
#pragma unroll
    for (int r = 0; r < 16; r++)
        y[r] = c_tab.hw[unfold_w(r)] * (unfold_sc(r) * oc[unfold_idx(r)] + unfold_ss(r) * os[unfold_idx(r)]);
}

// hardware sin/cos of pi*t (v_sin_f32 / v_cos_f32 take the angle in revolutions)
__device__ __forceinline__ void sincos_pi(float t, float* s, float* c) {
    *s = __builtin_amdgcn_sinf(0.5f * t);
    *c = __builtin_amdgcn_cosf(0.5f * t);
}

// cos/sin of (k + 0.5) * pi * t for k = 0..7 by rotation recurrence (2 sin/cos + 7 rotations).
__device__ __forceinline__ void half_phase8(float t, float* c, float* s) {
    float s1, c1, sh, ch;
    sincos_pi(t, &s1, &c1);
    sincos_pi(0.5f * t, &sh, &ch);
    c[0] = ch;
    s[0] = sh;
#pragma unroll
    for (int k = 1; k < 8; k++) {
        c[k] = fmaf(c[k - 1], c1, -s[k - 1] * s1);
        s[k] = fmaf(s[k - 1], c1, c[k - 1] * s1);
    }
}

// ---------------------------------------------------------------- inverse CLT of one unit
// LDS scratch of one 16-lane unit: 16 rows of 20 floats (a 16-byte row store of 8 consecutive
// lanes covers all 32 banks once), units 336 floats apart (neighbouring units use the other
// half of the banks when a wavefront reads them together).
constexpr int IMCLT_SCR_ROW = 20;
constexpr int IMCLT_SCR_UNIT = 336;

// Lane (hs = l >> 3, kh = l & 7) enters with p[kv] = CLT[hs][kh][kv], q[kv] = CLT[hs | 2][kh][kv];
// scratch: IMCLT_SCR_UNIT floats of LDS private to this unit.  Leaves with row[c] = output row
// r = l of the 16x16 pixel tile (sum over the 4 modes of unfold(IDTT), ImageDtt.java:3620-3632
// without the 1/4).
__device__ __forceinline__ void imclt_unit(int l, const float* p, const float* q, float* scratch, float* row) {
    const int hs = l >> 3, kh = l & 7;
    float oc[8], os[8], col[16];
    dct_iv8(p, oc);                          // vertical inverse of the C (v) modes
    dst_iv8(q, os);                          // vertical inverse of the S (v) modes
    unfold16(oc, os, col);                   // 16 rows for (hs, kh)
#pragma unroll
    for (int r = 0; r < 16; r += 4)
        *reinterpret_cast<float4*>(&scratch[(hs * 8 + kh) * IMCLT_SCR_ROW + r]) =
            make_float4(col[r], col[r + 1], col[r + 2], col[r + 3]);
    wave_sync();
    float rc[8], rs[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        rc[k] = scratch[k * IMCLT_SCR_ROW + l];
        rs[k] = scratch[(8 + k) * IMCLT_SCR_ROW + l];
    }
    wave_sync();
    dct_iv8(rc, oc);                         // horizontal inverse, C (h) modes
    dst_iv8(rs, os);                         // horizontal inverse, S (h) modes
    unfold16(oc, os, row);
}

}  // namespace clt
