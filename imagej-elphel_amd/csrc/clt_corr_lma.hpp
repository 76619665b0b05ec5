// Per-tile correlation LMA and foreground correction (gfx950).
//
// The reference refines every tile's disparity after the centre-of-mass step (ImageDtt.java
// ID:2254-2336): a Levenberg-Marquardt fit of a paraboloid to the combined pair correlations
// (Correlation2d.corrLMA, C2:1754-1900; sample selection filterConvex C2:2029-2207 and addSamples
// C2:2230-2338; the fit Correlations2dLMA.java), four one-direction fits (corr4dirsLMA,
// C2:1591-1630) and the foreground correction (foregroundCorrect, C2:1663-1750).
//
// Mapping.  Samples: 16 lanes (one DPP row) per tile, 4 tiles per wavefront; the fit over all
// pairs uses the tile's 16 lanes, the four one-direction fits run two at a time on 8 lanes each.
// A fit's candidate samples are the 9x9 box around its start point: the lanes own box cells fl,
// fl+W, ..., the combined correlation values sit in LDS as doubles, the convexity flags are
// gathered into an 81-bit mask with ballots and the selection grows by shifts of that mask (no
// memory round trips), the selected cells are compacted into a per-group sample list, and the
// fit's samples are reduced once to the moments its model needs (LmaState).  Iterations: one
// thread per fit, from the state alone.  Arithmetic is FP64 like the reference; the damped normal
// system is solved by LDL^T in registers.
#pragma once

namespace clt {

constexpr int LMA_TILES_PER_WAVE = 4;
constexpr int LMA_WAVES = 4;                 // 256-thread blocks
constexpr int LMA_SLOTS = LMA_TILES_PER_WAVE * LMA_WAVES;
constexpr int LMA_CELLS = 81;                // 9x9 box: half window at most 4
constexpr int LMA_MAXG = 2;                  // compatible pair groups of the fit over all pairs
constexpr int LMA_GPAIRS = 4;                // pairs per group (the 8-camera table has at most 4)
constexpr int LMA_DPAIRS = 2;                // pairs per group of a one-direction fit
constexpr int LMA_NFITS = 6;                 // all pairs, horizontal, vertical, diag main, diag other, one pair
constexpr int LMA_FIT_SINGLE = 5;            // a single correlation as a horizontal pair (C2:1903-1935)
constexpr int LMA_NMOM = 14;                 // sample moments of one group

struct LmaGroup {
    int32_t npairs, diag;
    double scale;
    int32_t pair[LMA_GPAIRS];
    int32_t xform[LMA_GPAIRS];               // 0 as is, 1 transposed (vertical), 2 upside down (other diagonal)
};

struct LmaFit {
    int32_t ngroups, reserved;
    LmaGroup g[LMA_MAXG];
};

struct LmaArgs {
    LmaFit fit[LMA_NFITS];
    double min_poly_strength, vasw_pwr, nc_cost, half_width, cost_wm, cost_wxy;
    double lambda0, l_good, l_bad, l_max, rms_diff;
    double fo_min_strength, fo_min_eff, fo_min_eff_ratio, fo_max_hwidth, fo_min_diff, fo_overcorr, fo_lim_overcorr;
    int32_t hwin, add3x3, adj_wm, adj_wyd, adj_wxy, adj_ag, num_iter, mix, fo_correct, fo_far, fo_ortho, npairs;
};

// ---- reductions over the W lanes of one fit (DPP; identical result in all W lanes) ---------------
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int W>
__device__ __forceinline__ double fit_sum(double v) {
    if constexpr (W == 16) {   // row_ror 8, 4, 2, 1
        v += dpp_f64<0x128>(v);
        v += dpp_f64<0x124>(v);
        v += dpp_f64<0x122>(v);
        v += dpp_f64<0x121>(v);
    } else {                   // quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror
        v += dpp_f64<0xb1>(v);
        v += dpp_f64<0x4e>(v);
        v += dpp_f64<0x141>(v);
    }
    return v;
}
template <int W>
__device__ __forceinline__ double fit_max(double v) {
    if constexpr (W == 16) {
        v = fmax(v, dpp_f64<0x128>(v));
        v = fmax(v, dpp_f64<0x124>(v));
        v = fmax(v, dpp_f64<0x122>(v));
        v = fmax(v, dpp_f64<0x121>(v));
    } else {
        v = fmax(v, dpp_f64<0xb1>(v));
        v = fmax(v, dpp_f64<0x4e>(v));
        v = fmax(v, dpp_f64<0x141>(v));
    }
    return v;
}

// 1/x to about 1 ulp: hardware estimate and two Newton steps
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// |v|^e for a value-as-weight power other than 2 (out of line: its constants would otherwise stay
// in registers through the whole fit)
__device__ __noinline__ double lma_pow(double v, double e) { return pow(fabs(v), e); }

// index of (i, j), i <= j, in a packed upper triangle of n x n
__host__ __device__ constexpr int lma_tri(int n, int i, int j) { return i * n - i * (i - 1) / 2 + (j - i); }

// ---- the 9x9 box as an 81-bit mask (bit br * 9 + bc) -------------------------------------------
struct Box81 {
    uint64_t lo, hi;   // bits 0..63, 64..80
};
__device__ __forceinline__ Box81 b_and(Box81 a, Box81 b) { return {a.lo & b.lo, a.hi & b.hi}; }
__device__ __forceinline__ Box81 b_or(Box81 a, Box81 b) { return {a.lo | b.lo, a.hi | b.hi}; }
__device__ __forceinline__ Box81 b_andn(Box81 a, Box81 b) { return {a.lo & ~b.lo, a.hi & ~b.hi}; }
__device__ __forceinline__ bool b_any(Box81 a) { return (a.lo | a.hi) != 0; }
template <int N> __device__ __forceinline__ Box81 b_shl(Box81 a) { return {a.lo << N, ((a.hi << N) | (a.lo >> (64 - N))) & 0x1ffffull}; }
template <int N> __device__ __forceinline__ Box81 b_shr(Box81 a) { return {(a.lo >> N) | (a.hi << (64 - N)), a.hi >> N}; }
__device__ __forceinline__ bool b_bit(Box81 a, int c) { return ((c < 64 ? a.lo >> c : a.hi >> (c - 64)) & 1) != 0; }
// number of set bits below bit c
__device__ __forceinline__ int b_rank(Box81 a, int c) {
    return c < 64 ? __popcll(a.lo & ((1ull << c) - 1)) : __popcll(a.lo) + __popcll(a.hi & ((1ull << (c - 64)) - 1));
}
__device__ __forceinline__ int b_count(Box81 a) { return __popcll(a.lo) + __popcll(a.hi); }
// bits [s, e) of a 64-bit word, 0 <= s <= e <= 64
__device__ __forceinline__ uint64_t bits64(int s, int e) {
    const uint64_t hi = (e >= 64) ? ~0ull : ((1ull << e) - 1);
    return (s >= e) ? 0ull : (hi & ~((1ull << s) - 1));
}
// rows a..b (empty when a > b): bits [9a, 9b + 9)
__device__ __forceinline__ Box81 b_rows(int a, int b) {
    if (a > b) return {0, 0};
    const int s = 9 * a, e = 9 * b + 9;
    return {bits64(min(s, 64), min(e, 64)), bits64(max(s, 64) - 64, max(e, 64) - 64)};
}
// columns a..b of every row (empty when a > b): the 9-bit row pattern times sum_r 2^(9 r)
__device__ __forceinline__ Box81 b_cols(int a, int b) {
    const uint64_t m9 = (a > b) ? 0 : bits64(a, b + 1);
    constexpr uint64_t COL0_LO = 0x8040201008040201ull;   // bits 0, 9, ..., 63
    return {COL0_LO * m9, __umul64hi(COL0_LO, m9) + (m9 << 8)};   // row 8 starts at bit 72 = 64 + 8
}

// ---- the same box as rows on the lanes of a fit (the ring growth of the selection) ------------
// Lane j = fl & 7 of a fit holds box row j in bits 0..8, lane j = 0 also row 8 in bits 16..24; a
// 16-lane fit keeps two identical copies (lanes 0..7 and 8..15).  A mask operation is one 32-bit
// operation and a move by a row is a DPP row shift.  Every lane of the 16-lane DPP row must be
// active when r9_up, r9_down or r9_to_box run, or at least every lane of the caller's group.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t r9_rows(int j, int a, int b) {   // rows a..b (empty when a > b)
    uint32_t m = (j >= a && j <= b) ? 0x1ffu : 0u;
    if (j == 0 && a <= 8 && b >= 8) m |= 0x1ff0000u;
    return m;
}
__device__ __forceinline__ uint32_t r9_cols(int a, int b) {          // columns a..b of every row
    const uint32_t m = (a > b) ? 0u : (((2u << b) - 1u) & ~((1u << a) - 1u));
    return m | (m << 16);
}
// row r + 1 -> row r (the linear mask shifted right by 9)
__device__ __forceinline__ uint32_t r9_up(uint32_t x, int j) {
    const uint32_t nxt = dpp_u32<0x101>(x);   // row_shl:1, lane + 1
    const uint32_t r8 = dpp_u32<0x117>(x);    // row_shr:7, lane j = 7 reads lane j = 0 of its group
    return (j == 7) ? (r8 >> 16) : (nxt & 0x1ffu);
}
// row r -> row r + 1 (shifted left by 9; row 8 leaves the box)
__device__ __forceinline__ uint32_t r9_down(uint32_t x, int j) {
    const uint32_t prv = dpp_u32<0x111>(x);   // row_shr:1, lane - 1
    const uint32_t r7 = dpp_u32<0x107>(x);    // row_shl:7, lane j = 0 reads lane j = 7 of its group
    return (j == 0) ? ((r7 & 0x1ffu) << 16) : (prv & 0x1ffu);
}
__device__ __forceinline__ uint32_t r9_from_box(Box81 a, int j) {
    uint32_t r = (j < 7) ? ((uint32_t)(a.lo >> (9 * j)) & 0x1ffu)
                         : (((uint32_t)(a.lo >> 63) | ((uint32_t)a.hi << 1)) & 0x1ffu);   // bits 63..71
    if (j == 0) r |= ((uint32_t)(a.hi >> 8) & 0x1ffu) << 16;                              // bits 72..80
    return r;
}
// the linear mask in every lane: OR of the 8 rows of the group (quad_perm [1,0,3,2], [2,3,0,1],
// row_half_mirror)
__device__ __forceinline__ Box81 r9_to_box(uint32_t x, int j) {
    const uint64_t lo = (uint64_t)(x & 0x1ffu) << (9 * j);   // row 7 keeps only bit 63 here
    uint32_t l0 = (uint32_t)lo, l1 = (uint32_t)(lo >> 32);
    uint32_t h = (j == 7) ? ((x & 0x1ffu) >> 1) : 0u;
    if (j == 0) h |= ((x >> 16) & 0x1ffu) << 8;
    l0 |= dpp_u32<0xb1>(l0); l1 |= dpp_u32<0xb1>(l1); h |= dpp_u32<0xb1>(h);
    l0 |= dpp_u32<0x4e>(l0); l1 |= dpp_u32<0x4e>(l1); h |= dpp_u32<0x4e>(h);
    l0 |= dpp_u32<0x141>(l0); l1 |= dpp_u32<0x141>(l1); h |= dpp_u32<0x141>(h);
    return {((uint64_t)l1 << 32) | l0, (uint64_t)h};
}

struct LmaOut {
    bool ok;
    double disp, strength, wm, wxy;
};

// What the iterations of one fit need from its samples (written by lma_fit, read by lma_iterate):
// the sample moments of each group about xo (the box centre on the group's disparity axis), the
// group constants and the start.  The model is a paraboloid in the sample coordinates: with
// u = x - X0, Y = y^2, alpha = ws^2 / Wx^2, beta = ws^2 / Wy^2 its shape is d = 1 - alpha u^2 -
// beta Y and every derivative is a u, u^2 or Y term, so the weighted residual square sum and the
// normal equations at any parameter vector are polynomials in the sums of sw x'^i (i <= 4),
// sw x'^i Y (i <= 2), sw Y^2, sw v x'^i (i <= 2), sw v Y and sw v^2 with x' = x - xo: the same
// sums as the per-sample passes of getFxJt / getWJtJlambda / getWYmFxRms
// (Correlations2dLMA:368-461), reassociated.
struct LmaState {
    double mom[LMA_MAXG][LMA_NMOM];
    double xo[LMA_MAXG], scale[LMA_MAXG], ws2[LMA_MAXG], gwn[LMA_MAXG];
    double x0, vmax, wr0, wr1;
    int32_t groups;   // bit g: group g has samples
    int32_t run;      // 1: the fit has samples (lma_iterate runs it), 0: it fails
};

// Per-fit LDS of one tile: values of the box cells (group g at +81 g), compacted sample lists and
// the fit's state for the iterations.
struct LmaLds {
    double* v;
    uint16_t* idx;   // cell | weight class << 8 (1 convex, 2 other selected, 0 NaN sample)
    LmaState* st;
};

// LDL^T of a symmetric n x n matrix (packed upper triangle a, constant indices only) and the
// solution of a x = b; returns the product of the pivots (the determinant).  A zero pivot leaves x
// undefined.
template <int N>
__device__ __forceinline__ double ldl_solve(const double* a, const double* b, double* x) {
    double L[N][N], d[N], det = 1.0;
#pragma unroll
    for (int k = 0; k < N; k++) {
        double dk = a[lma_tri(N, k, k)];
#pragma unroll
        for (int s = 0; s < k; s++) dk -= L[k][s] * L[k][s] * d[s];
        d[k] = dk;
        det *= dk;
        const double idk = 1.0 / dk;
#pragma unroll
        for (int i = k + 1; i < N; i++) {
            double m = a[lma_tri(N, k, i)];
#pragma unroll
            for (int s = 0; s < k; s++) m -= L[i][s] * L[k][s] * d[s];
            L[i][k] = m * idk;
        }
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
        double m = b[i];
#pragma unroll
        for (int s = 0; s < i; s++) m -= L[i][s] * x[s];
        x[i] = m;
    }
#pragma unroll
    for (int i = N - 1; i >= 0; i--) {
        double m = x[i] / d[i];
#pragma unroll
        for (int s = i + 1; s < N; s++) m -= L[s][i] * x[s];
        x[i] = m;
    }
    return det;
}

// product of the row norms of a symmetric matrix given as a packed upper triangle (PA:590-598)
template <int N>
__device__ __forceinline__ double tri_norm(const double* a) {
    double norm = 1.0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        double r = 0.0;
#pragma unroll
        for (int j = 0; j < N; j++) {
            const double v = a[i <= j ? lma_tri(N, i, j) : lma_tri(N, j, i)];
            r += v * v;
        }
        norm *= sqrt(r);
    }
    return norm;
}

// The samples of one Correlations2dLMA fit (C2:1754-1900) over the groups of *F, started at
// xcenter (pixels along the disparity axis, = -disparity), on W lanes (fl = lane within the fit,
// fmask = the fit's lanes in the wavefront): the box, its convex cells, the selection, the weights
// and the sample moments, left in *S.st for lma_iterate (run = 0 when the fit has no samples).
// Groups of at most NPM pairs.  The tile's correlations start at corr[toff] (a wave-uniform base
// and 32-bit offsets keep one register per load address).  POLY: instead, the 2-D quadratic
// maximum of the same samples (run_poly_instead, getMaxXYPoly Correlations2dLMA:643-705),
// returned: disp = x, wm = y, strength = the value.
template <int W, int NC, int NPM, bool POLY = false>
__device__ __forceinline__ LmaOut lma_fit(const LmaArgs& A, const LmaFit* F, const float* __restrict__ corr, int toff,
                                          double xcenter, int fl, int lane, uint64_t fmask, const LmaLds& S) {
    constexpr int CPL = (LMA_CELLS + W - 1) / W;   // box cells per lane
    LmaOut out = {false, 0.0, 0.0, 0.0, 0.0};
    if (fl == 0) S.st->run = 0;
    const int G = F->ngroups;
    if (G == 0) return out;
    const int hwin = A.hwin;
    int ns[LMA_MAXG] = {0, 0}, x0s[LMA_MAXG] = {0, 0}, y0s[LMA_MAXG] = {0, 0};
    double gw_raw[LMA_MAXG] = {0.0, 0.0}, sw_fit = 0.0, vmax = -INFINITY;
    Box81 cvs[LMA_MAXG] = {{0, 0}, {0, 0}};   // selected convex cells (weight class 1)
#pragma unroll
    for (int g = 0; g < LMA_MAXG; g++) {
        if (g >= G) break;
        const LmaGroup& grp = F->g[g];
        const int npairs = grp.npairs;
        const int ixc = (int)floor(xcenter / grp.scale + 0.5);   // Math.round (C2:2256)
        const int x0 = ixc + 7, y0 = grp.diag ? ixc + 7 : 7;
        x0s[g] = x0;
        y0s[g] = y0;
        double* v = S.v + g * LMA_CELLS;
        uint16_t* idx = S.idx + g * LMA_CELLS;
        // combined correlation of the group's pairs over the box (combineCompatiblePairs, C2:462-488):
        // every load of the lane's cells is issued before the first is used
        // element (row, col) of pair q's correlation as seen by the group: base + a row + b col
        // (as is, transposed for a vertical pair, upside down for the other diagonal)
        int base[NPM], ra[NPM], cb[NPM];
#pragma unroll
        for (int q = 0; q < NPM; q++) {
            const int xf = grp.xform[q < npairs ? q : 0];
            base[q] = toff + grp.pair[q < npairs ? q : 0] * CLT_CORR_LEN + ((xf == 2) ? 210 : 0);
            ra[q] = (xf == 1) ? 1 : ((xf == 2) ? -15 : 15);
            cb[q] = (xf == 1) ? 15 : 1;
        }
        float raw[CPL][NPM];
        uint32_t in_grid = 0;   // bit k: cell fl + W k lies on the 15 x 15 correlation
        int r0s[CPL], c0s[CPL];
#pragma unroll
        for (int k = 0; k < CPL; k++) {
            const int c = fl + W * k;
            const int br = (c * 57) >> 9, bc = c - 9 * br;   // c / 9, c % 9 for c < 81
            const int row = y0 - 4 + br, col = x0 - 4 + bc;
            const bool ing = c < LMA_CELLS && row >= 0 && row < 15 && col >= 0 && col < 15;
            in_grid |= ing ? (1u << k) : 0u;
            r0s[k] = ing ? row : 0;   // loads stay unconditional and in bounds
            c0s[k] = ing ? col : 0;
        }
        // the group's pairs only (a wave-uniform count); each load a 32-bit byte offset from the
        // wave-uniform base (|strides| <= 15, r0, c0 < 15: 24-bit products are exact)
        const char* cbase = reinterpret_cast<const char*>(corr);
        auto load_box = [&](auto nq) {
            constexpr int NQ = decltype(nq)::value;
#pragma unroll
            for (int k = 0; k < CPL; k++)
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const uint32_t off = (uint32_t)(base[q] + __mul24(ra[q], r0s[k]) + __mul24(cb[q], c0s[k])) * 4u;
                    raw[k][q] = *reinterpret_cast<const float*>(cbase + off);
                }
        };
        if (npairs >= NPM) load_box(std::integral_constant<int, NPM>{});
        else if (npairs == 1) load_box(std::integral_constant<int, 1>{});
        else if (npairs == 2 && NPM > 2) load_box(std::integral_constant<int, (NPM > 2 ? 2 : NPM)>{});
        else load_box(std::integral_constant<int, NPM>{});   // (loads of pairs past npairs are discarded)
        __asm__ volatile("" ::: "memory");   // all loads in flight before the first use
#pragma unroll
        for (int k = 0; k < CPL; k++) {
            const bool ing = (in_grid >> k) & 1;
#pragma unroll
            for (int q = 0; q < NPM; q++) raw[k][q] = (ing && q < npairs) ? raw[k][q] : 0.f;
        }
        // the reference divides the sum by the pair count; for 2 and 4 a product is the same number
        if ((npairs & (npairs - 1)) == 0) {   // n = 1, 2, 4: the product by 1/n is the division exactly
            const double rn = (npairs == 4) ? 0.25 : ((npairs == 2) ? 0.5 : 1.0);
#pragma unroll
            for (int k = 0; k < CPL; k++) {
                const int c = fl + W * k;
                double val = 0.0;
#pragma unroll
                for (int q = 0; q < NPM; q++) val += (double)raw[k][q];
                if (c < LMA_CELLS) v[c] = val * rn;
            }
        } else {
#pragma unroll
            for (int k = 0; k < CPL; k++) {
                const int c = fl + W * k;
                double val = 0.0;
#pragma unroll
                for (int q = 0; q < NPM; q++) val += (double)raw[k][q];
                if (c < LMA_CELLS) v[c] = val / npairs;
            }
        }
        wave_sync();
        // bi-convex cells inside the clamped window (C2:2046-2064), in box coordinates (box row br is
        // frame row y0 - 4 + br); the 81-bit masks are the same in all lanes of the fit
        const int gr0 = max(4 - y0, 0), gr1 = min(18 - y0, 8), gc0 = max(4 - x0, 0), gc1 = min(18 - x0, 8);
        const int wr0 = max(4 - hwin, gr0), wr1 = min(4 + hwin, gr1), wc0 = max(4 - hwin, gc0), wc1 = min(4 + hwin, gc1);
        Box81 cv = {0, 0};
#pragma unroll 1
        for (int k = 0; k < CPL; k++) {
            const int c = fl + W * k;
            const int br = (c * 57) >> 9, bc = c - 9 * br;
            bool conv = false;
            if (c < LMA_CELLS && br > wr0 && br < wr1 && bc > wc0 && bc < wc1) {
                const double d = v[c];
                conv = (2 * d > (v[c - 1] + v[c + 1])) && (2 * d > (v[c - 9] + v[c + 9]));
            }
            const uint64_t bits = (__ballot(conv) & fmask) >> (lane - fl);   // cells W k .. W k + W - 1
            if (W * k < 64) cv.lo |= bits << (W * k);
            if (W * k + W > 64) cv.hi |= (W * k >= 64) ? (bits << (W * k - 64)) : (bits >> (64 - W * k));
        }
        cv.hi &= 0x1ffffull;
        // the selection grows ring by ring from the start cell (C2:2066-2195): the sources of step
        // istep are the selected cells of ring istep, convex ones (any on the first step).  On the
        // rows form; the four moves are masked, not branched, so every lane takes part in the DPP
        const int j = fl & 7;
        const uint32_t cvr = r9_from_box(cv, j);
        uint32_t sel = (4 >= wr0 && 4 <= wr1 && 4 >= wc0 && 4 <= wc1) ? (r9_rows(j, 4, 4) & r9_cols(4, 4)) : 0u;
        const uint32_t dl = r9_cols(wc0 + 1, 4), dr = r9_cols(4, wc1 - 1);         // diagonal steps of vertical moves
        const uint32_t du = r9_rows(j, wr0 + 1, 4), dd = r9_rows(j, 4, wr1 - 1);   // ... of horizontal moves
        for (int istep = 0; istep < hwin; istep++) {
            const uint32_t src = (istep == 0) ? sel : (sel & cvr);
            const uint32_t scols = r9_cols(max(4 - istep, gc0), min(4 + istep, gc1));
            const uint32_t srows = r9_rows(j, max(4 - istep, gr0), min(4 + istep, gr1));
            const uint32_t m_up = (istep < y0) ? ~0u : 0u, m_dn = (4 + istep < wr1) ? ~0u : 0u;
            const uint32_t m_lf = (istep < x0) ? ~0u : 0u, m_rt = (4 + istep < wc1) ? ~0u : 0u;
            const uint32_t tu = r9_up(src & r9_rows(j, 4 - istep, 4 - istep) & scols & m_up, j);
            const uint32_t td = r9_down(src & r9_rows(j, 4 + istep, 4 + istep) & scols & m_dn, j);
            const uint32_t tl = (src & r9_cols(4 - istep, 4 - istep) & srows & m_lf) >> 1;
            const uint32_t tr = (src & r9_cols(4 + istep, 4 + istep) & srows & m_rt) << 1;
            const uint32_t dv = (tu | td) & cvr, dh = (tl | tr) & cvr;
            const uint32_t nw = sel | tu | td | tl | tr | ((dv & dl) >> 1) | ((dv & dr) << 1) | r9_up(dh & du, j) |
                                r9_down(dh & dd, j);
            const bool grew = (__ballot((nw & ~sel) != 0u) & fmask) != 0;
            sel = nw;
            if (!grew) break;
        }
        const Box81 selb = r9_to_box(sel, j);
        // cell weights (C2:2197-2206): selected convex 1, other selected nc_cost, with add3x3 the
        // 3x3 around the start at least nc_cost; the samples (addSamples, C2:2230-2338) in cell
        // order; NaN values or weights enter the fit as 0 (Correlations2dLMA:262-266)
        const Box81 grid = b_and(b_rows(gr0, gr1), b_cols(gc0, gc1));
        Box81 smp = selb;
        if (A.add3x3) smp = b_or(smp, b_and(b_rows(3, 5), b_cols(3, 5)));
        smp = b_and(smp, grid);
        if (!(A.nc_cost > 0.0)) smp = b_and(selb, cv);   // the non-convex cells weigh nothing
        cvs[g] = b_and(selb, cv);                      // weight class 1
        double gsum = 0.0;
#pragma unroll 1
        for (int k = 0; k < CPL; k++) {
            const int c = fl + W * k;
            bool nan = false;
            if (c < LMA_CELLS && b_bit(smp, c)) {
                const double val = v[c];
                double w = (b_bit(cvs[g], c) ? 1.0 : A.nc_cost) * npairs;
                if (A.vasw_pwr != 0.0) w *= (A.vasw_pwr == 2.0) ? val * val : lma_pow(val, A.vasw_pwr);
                gsum += w;
                nan = isnan(val) || isnan(w);
                if (nan) v[c] = 0.0;
                sw_fit += nan ? 0.0 : w;
                vmax = fmax(vmax, nan ? 0.0 : val);
                const int cls = nan ? 0 : (b_bit(cvs[g], c) ? 1 : 2);
                idx[b_rank(smp, c)] = (uint16_t)(c | cls << 8);
            }
        }
        ns[g] = b_count(smp);
        gw_raw[g] = fit_sum<W>(gsum);
    }
    // groups are numbered in the order of their first sample (Correlations2dLMA:196-210)
    const bool ex0 = ns[0] > 0, ex1 = G > 1 && ns[1] > 0;
    if (!ex0 && !ex1) return out;
    sw_fit = fit_sum<W>(sw_fit);
    vmax = fit_max<W>(vmax);
    // setWeightsValues (Correlations2dLMA:245-278): unit total weight with the two regularisation
    // points (Wm -> half_width at cost_wm, Wxy -> 0 at cost_wxy)
    const double tot = sw_fit + A.cost_wm + A.cost_wxy;
    const double inv = (tot != 0.0) ? 1.0 / tot : 1.0;
    wave_sync();
    const double sqrt2 = 1.4142135623730951;
    double wc[LMA_MAXG][2];   // weight of a class-1 / class-2 sample before |v|^pwr, normalised
#pragma unroll
    for (int g = 0; g < LMA_MAXG; g++) {
        const int np = (g < G) ? F->g[g].npairs : 0;
        wc[g][0] = (1.0 * np) * inv;
        wc[g][1] = (A.nc_cost * np) * inv;
    }

    // sample j of group g: value, weight and coordinates (x along the group's disparity axis)
    auto sample = [&](int g, int j, double& sv, double& sw, double& x, double& y) {
        const LmaGroup& grp = F->g[g];
        const int e = S.idx[g * LMA_CELLS + j];
        const int c = e & 0xff, cls = e >> 8;
        sv = S.v[g * LMA_CELLS + c];
        sw = (cls == 1) ? wc[g][0] : ((cls == 2) ? wc[g][1] : 0.0);
        if (A.vasw_pwr != 0.0) sw *= (A.vasw_pwr == 2.0) ? sv * sv : lma_pow(sv, A.vasw_pwr);
        const int br = (c * 57) >> 9, bc = c - 9 * br;
        const int cx = x0s[g] - 11 + bc, cy = y0s[g] - 11 + br;   // relative to the correlation centre
        const double hx = grp.diag ? 0.5 * grp.scale : grp.scale;
        x = (double)(grp.diag ? cx + cy : cx) * hx;
        y = (double)(grp.diag ? cy - cx : cy) * hx;
    };
    if constexpr (POLY) {
        // PolynomialApproximation.quadraticApproximation (PA:350-580) and quadraticMaxV2dX2Y2XY
        // (PA:201-224): weighted least squares z ~ A x^2 + B y^2 + C xy + D x + E y + F over the
        // samples with w > 0 (the weights' common scale cancels in the solution and in the
        // determinant / row-norm tests), then the stationary point
        double m[21];
#pragma unroll
        for (int i = 0; i < 21; i++) m[i] = 0.0;
#pragma unroll
        for (int g = 0; g < LMA_MAXG; g++) {
            if (g >= G || ns[g] == 0) continue;
            for (int j = fl; j < ns[g]; j += W) {
                double sv, sw, x, y;
                sample(g, j, sv, sw, x, y);
                if (!(sw > 0.0)) continue;
                const double x2 = x * x, y2 = y * y, wz = sw * sv;
                m[0] += sw;            m[1] += sw * x;        m[2] += sw * y;
                m[3] += sw * x2;       m[4] += sw * x * y;    m[5] += sw * y2;
                m[6] += sw * x2 * x;   m[7] += sw * x2 * y;   m[8] += sw * x * y2;   m[9] += sw * y2 * y;
                m[10] += sw * x2 * x2; m[11] += sw * x2 * x * y; m[12] += sw * x2 * y2;
                m[13] += sw * x * y2 * y; m[14] += sw * y2 * y2;
                m[15] += wz;           m[16] += wz * x;       m[17] += wz * y;
                m[18] += wz * x2;      m[19] += wz * x * y;   m[20] += wz * y2;
            }
        }
#pragma unroll
        for (int i = 0; i < 21; i++) m[i] = fit_sum<W>(m[i]);
        const double S00 = m[0], S10 = m[1], S01 = m[2], S20 = m[3], S11 = m[4], S02 = m[5], S30 = m[6], S21 = m[7];
        const double S12 = m[8], S03 = m[9], S40 = m[10], S31 = m[11], S22 = m[12], S13 = m[13], S04 = m[14];
        const double lin[6] = {S20, S11, S10, S02, S01, S00};   // packed upper triangle, 3 x 3
        const double zl[3] = {0.0, 0.0, 0.0};
        double xl[3];
        const double nml = tri_norm<3>(lin);
        const double detl = ldl_solve<3>(lin, zl, xl);
        if (nml == 0.0 || fabs(detl) / nml < 1e-20) return out;   // linear only: null
        const double quad[21] = {S40, S22, S31, S30, S21, S20, S04, S13, S12, S03, S02, S22, S21, S12, S11,
                                 S20, S11, S10, S02, S01, S00};
        const double zq[6] = {m[18], m[20], m[19], m[16], m[17], m[15]};
        double c[6];
        const double nmq = tri_norm<6>(quad);
        const double detq = ldl_solve<6>(quad, zq, c);
        if (nmq == 0.0 || fabs(detq) / nmq < 1e-30) return out;
        const double m00 = 2 * c[0], m01 = c[2], m11 = 2 * c[1];
        const double det2 = m00 * m11 - m01 * m01;
        const double nm2 = sqrt(m00 * m00 + m01 * m01) * sqrt(m01 * m01 + m11 * m11);
        if (nm2 == 0.0 || fabs(det2) / nm2 < 1e-30) return out;
        const double px = (-c[3] * m11 + c[4] * m01) / det2, py = (-c[4] * m00 + c[3] * m01) / det2;
        const double v = c[0] * px * px + c[1] * py * py + c[2] * px * py + c[3] * px + c[4] * py + c[5];
        if (v < 0.0 || c[0] >= 0.0 || c[1] >= 0.0) return out;   // getMaxXYPoly: a maximum only
        out.ok = true;
        out.disp = px;
        out.wm = py;
        out.strength = v;
        return out;
    }
    // sample moments of each group about the box centre, and the fit's constants
    LmaState* st = S.st;
#pragma unroll
    for (int g = 0; g < LMA_MAXG; g++) {
        if (g >= G || ns[g] == 0) continue;
        const LmaGroup& grp = F->g[g];
        const int cx = x0s[g] - 7, cy = y0s[g] - 7;   // the box centre cell
        const double xo = (double)(grp.diag ? cx + cy : cx) * (grp.diag ? 0.5 * grp.scale : grp.scale);
        double m[LMA_NMOM];
#pragma unroll
        for (int k = 0; k < LMA_NMOM; k++) m[k] = 0.0;
        for (int j = fl; j < ns[g]; j += W) {
            double sv, sw, x, y;
            sample(g, j, sv, sw, x, y);
            const double xp = x - xo, Y = y * y;
            const double w1 = sw * xp, w2 = w1 * xp, wy = sw * Y, wv = sw * sv, wv1 = wv * xp;
            m[0] += sw;        m[1] += w1;        m[2] += w2;        m[3] += w2 * xp;   m[4] += w2 * (xp * xp);
            m[5] += wy;        m[6] += wy * xp;   m[7] += wy * (xp * xp);               m[8] += wy * Y;
            m[9] += wv;        m[10] += wv1;      m[11] += wv1 * xp; m[12] += wv * Y;   m[13] += wv * sv;
        }
#pragma unroll
        for (int k = 0; k < LMA_NMOM; k++) m[k] = fit_sum<W>(m[k]);
        if (fl == 0) {
#pragma unroll
            for (int k = 0; k < LMA_NMOM; k++) st->mom[g][k] = m[k];
            st->xo[g] = xo;
            st->scale[g] = grp.scale;
            const double ws = grp.diag ? sqrt2 : 1.0;
            st->ws2[g] = ws * ws;
        }
    }
    if (fl == 0) {
#pragma unroll
        for (int g = 0; g < LMA_MAXG; g++) st->gwn[g] = (sw_fit > 0.0) ? gw_raw[g] / sw_fit : gw_raw[g];
        st->x0 = xcenter;
        st->vmax = vmax;
        st->wr0 = A.cost_wm * inv;
        st->wr1 = A.cost_wxy * inv;
        st->groups = (ex0 ? 1 : 0) | (ex1 ? 2 : 0);
        st->run = 1;
    }
    return out;
}

// The Levenberg-Marquardt iterations of one fit from its state (lmaStep and runLma,
// Correlations2dLMA:477-640; initVector :212-243; getDisparityStrength :165-185), all in this
// lane.  NC core parameters: X0, Wm, Wxy (3) or X0, Wm, Wyd, Wxy (4, when Wyd is adjusted); then
// one amplitude per group with samples.
template <int NC, bool IN_REGS = true>
__device__ __forceinline__ LmaOut lma_iterate(const LmaArgs& A, const LmaState* __restrict__ st_in) {
    // IN_REGS: the state (in device memory) copied to registers once for the whole fit; else read
    // where it is (LDS)
    LmaState stv;
    if constexpr (IN_REGS) stv = *st_in;
    const LmaState* st = IN_REGS ? &stv : st_in;
    constexpr int NP = NC + 2;
    // sums: core triangle, per group {core x amplitude, amplitude^2, J^T r of the amplitude},
    // J^T r of the core
    constexpr int I_CROSS = NC * (NC + 1) / 2, I_DIAG = I_CROSS + 2 * NC, I_BAG = I_DIAG + 2;
    constexpr int I_BC = I_BAG + 2, I_RMS = I_BC + NC;
    constexpr int P_WXY = NC - 1;
    LmaOut out = {false, 0.0, 0.0, 0.0, 0.0};
    const int groups = st->groups;
    const bool ex0 = groups & 1, ex1 = (groups >> 1) & 1;
    const double wr0 = st->wr0, wr1 = st->wr1, vmax = st->vmax;
    // parameters (initVector): X0, Wm, [Wyd,] Wxy, Ag of each group
    bool msk[NP];
    double p[NP], pn[NP];
    msk[0] = true;
    msk[1] = A.adj_wm != 0;
    if constexpr (NC == 4) msk[2] = A.adj_wyd != 0;
    msk[P_WXY] = A.adj_wxy != 0;
    msk[NC] = A.adj_ag != 0;
    msk[NC + 1] = A.adj_ag != 0 && ex0 && ex1;
#pragma unroll
    for (int i = 0; i < NP; i++) p[i] = 0.0;
    p[0] = st->x0;
    p[1] = A.half_width;
    p[NC] = p[NC + 1] = vmax;
    double acc[I_RMS];
    // the moments about X0 (binomial shift by X0 - xo)
    struct UMom {
        double s0, s1, s2, s3, s4, t0, t1, t2, q, v0, v1, v2, vy, vv;
    };
    auto umom = [&](int g, double x0) {
        const double* m = st->mom[g];
        const double d = x0 - st->xo[g];
        UMom u;
        u.s0 = m[0];
        u.s1 = m[1] - d * m[0];
        u.s2 = m[2] - d * (2.0 * m[1] - d * m[0]);
        u.s3 = m[3] - d * (3.0 * m[2] - d * (3.0 * m[1] - d * m[0]));
        u.s4 = m[4] - d * (4.0 * m[3] - d * (6.0 * m[2] - d * (4.0 * m[1] - d * m[0])));
        u.t0 = m[5];
        u.t1 = m[6] - d * m[5];
        u.t2 = m[7] - d * (2.0 * m[6] - d * m[5]);
        u.q = m[8];
        u.v0 = m[9];
        u.v1 = m[10] - d * m[9];
        u.v2 = m[11] - d * (2.0 * m[10] - d * m[9]);
        u.vy = m[12];
        u.vv = m[13];
        return u;
    };
    // `rms_at` the weighted residual square sum, `normal_at` the normal matrix and J^T W (y - f).
    // The LMA needs the normal equations only at accepted points it continues from.
    auto rms_at = [&](const double* q) {
        double s2 = 0.0;
#pragma unroll
        for (int g = 0; g < LMA_MAXG; g++) {
            if (!((groups >> g) & 1)) continue;
            const int gs = (g == 1 && ex0) ? 1 : 0;
            const double Wy = q[1] * st->scale[g] + ((NC == 4) ? q[2] : 0.0), Wx = Wy + q[P_WXY];
            const double iWx = rcp64(Wx), iWy = rcp64(Wy);
            const double Ag = gs ? q[NC + 1] : q[NC];
            const double ws2 = st->ws2[g], a = ws2 * iWx * iWx, b = ws2 * iWy * iWy;
            const UMom u = umom(g, q[0]);
            const double sd2 = u.s0 - 2.0 * (a * u.s2 + b * u.t0) + a * (a * u.s4 + 2.0 * b * u.t2) + b * b * u.q;
            const double sdv = u.v0 - a * u.v2 - b * u.vy;
            s2 += u.vv - Ag * (2.0 * sdv - Ag * sd2);
        }
        const double r0 = A.half_width - q[1], r1 = -q[P_WXY];
        return fmax(s2, 0.0) + wr0 * r0 * r0 + wr1 * r1 * r1;
    };
    auto normal_at = [&](const double* q) {
#pragma unroll
        for (int i = 0; i < I_RMS; i++) acc[i] = 0.0;
#pragma unroll
        for (int g = 0; g < LMA_MAXG; g++) {
            if (!((groups >> g) & 1)) continue;
            const int gs = (g == 1 && ex0) ? 1 : 0;   // amplitude slot of this group
            const double sc = st->scale[g];
            const double Wy = q[1] * sc + ((NC == 4) ? q[2] : 0.0), Wx = Wy + q[P_WXY];
            const double iWx = rcp64(Wx), iWy = rcp64(Wy);
            const double Ag = gs ? q[NC + 1] : q[NC];
            const double ws2 = st->ws2[g], a = ws2 * iWx * iWx, b = ws2 * iWy * iWy;
            const UMom u = umom(g, q[0]);
            // derivatives: X0 c0 u, Wm sc (cx u^2 + cy Y), Wyd cx u^2 + cy Y, Wxy cx u^2
            const double c0 = 2.0 * Ag * a, cx = c0 * iWx, cy = 2.0 * Ag * b * iWy;
            const double e = cx * (cx * u.s4 + 2.0 * cy * u.t2) + cy * cy * u.q;   // sum sw (cx u^2 + cy Y)^2
            const double f = cx * u.s3 + cy * u.t1;                               // sum sw u (cx u^2 + cy Y)
            const double h = cx * u.s4 + cy * u.t2;                               // sum sw u^2 (cx u^2 + cy Y)
            const double du = u.s1 - a * u.s3 - b * u.t1;                         // sum sw d u
            const double du2 = u.s2 - a * u.s4 - b * u.t2;                        // sum sw d u^2
            const double dy = u.t0 - a * u.t2 - b * u.q;                          // sum sw d Y
            const double sd2 = u.s0 - 2.0 * (a * u.s2 + b * u.t0) + a * (a * u.s4 + 2.0 * b * u.t2) + b * b * u.q;
            const double sdv = u.v0 - a * u.v2 - b * u.vy;
            double jj[NC][NC], cr[NC], jv[NC];
            jj[0][0] = c0 * c0 * u.s2;
            jj[0][1] = c0 * sc * f;
            jj[1][1] = sc * sc * e;
            jj[0][P_WXY] = c0 * cx * u.s3;
            jj[1][P_WXY] = sc * cx * h;
            jj[P_WXY][P_WXY] = cx * cx * u.s4;
            cr[0] = c0 * du;
            cr[1] = sc * (cx * du2 + cy * dy);
            cr[P_WXY] = cx * du2;
            jv[0] = c0 * u.v1;
            jv[1] = sc * (cx * u.v2 + cy * u.vy);
            jv[P_WXY] = cx * u.v2;
            if constexpr (NC == 4) {
                jj[0][2] = c0 * f;
                jj[1][2] = sc * e;
                jj[2][2] = e;
                jj[2][P_WXY] = cx * h;
                cr[2] = cx * du2 + cy * dy;
                jv[2] = cx * u.v2 + cy * u.vy;
            }
#pragma unroll
            for (int i = 0; i < NC; i++) {
#pragma unroll
                for (int k = i; k < NC; k++) acc[lma_tri(NC, i, k)] += jj[i][k];
                acc[I_CROSS + gs * NC + i] += cr[i];
                acc[I_BC + i] += jv[i] - Ag * cr[i];
            }
            acc[I_DIAG + gs] += sd2;
            acc[I_BAG + gs] += sdv - Ag * sd2;
        }
        // the two regularisation points
        const double r0 = A.half_width - q[1], r1 = -q[P_WXY];
        acc[lma_tri(NC, 1, 1)] += wr0;
        acc[lma_tri(NC, P_WXY, P_WXY)] += wr1;
        acc[I_BC + 1] += wr0 * r0;
        acc[I_BC + P_WXY] += wr1 * r1;
    };

    const double rms0 = sqrt(rms_at(p));
    double rms = rms0, lambda = A.lambda0;
    normal_at(p);
    bool good = false;
    for (int iter = 0; iter < A.num_iter; iter++) {
        // (J^T W J + lambda diag(J^T W J)) delta = J^T W (y - f) over the adjusted parameters
        // (lmaStep, Correlations2dLMA:533-640), LDL^T; a zero pivot is the singular matrix
        double L[NP][NP], dd[NP], x[NP];
        bool singular = false;
#pragma unroll
        for (int k = 0; k < NP; k++) {
            const double v0 = (k < NC) ? acc[lma_tri(NC, k, k)] : acc[I_DIAG + k - NC];
            double dk = msk[k] ? v0 + v0 * lambda : 1.0;
#pragma unroll
            for (int s = 0; s < k; s++) dk -= L[k][s] * L[k][s] * dd[s];
            dd[k] = dk;
            singular |= (dk == 0.0);
            const double idk = rcp64(dk);
#pragma unroll
            for (int i = k + 1; i < NP; i++) {
                double m = 0.0;
                if (msk[i] && msk[k]) {
                    if (i < NC) m = acc[lma_tri(NC, k, i)];
                    else if (k < NC) m = acc[I_CROSS + (i - NC) * NC + k];
                }
#pragma unroll
                for (int s = 0; s < k; s++) m -= L[i][s] * L[k][s] * dd[s];
                L[i][k] = m * idk;
            }
        }
        if (singular) break;   // lmaStep returns {false, true}
#pragma unroll
        for (int i = 0; i < NP; i++) {
            double m = msk[i] ? ((i < NC) ? acc[I_BC + i] : acc[I_BAG + i - NC]) : 0.0;
#pragma unroll
            for (int s = 0; s < i; s++) m -= L[i][s] * x[s];
            x[i] = m;
        }
#pragma unroll
        for (int i = NP - 1; i >= 0; i--) {
            double m = x[i] * rcp64(dd[i]);
#pragma unroll
            for (int s = i + 1; s < NP; s++) m -= L[s][i] * x[s];
            x[i] = m;
        }
#pragma unroll
        for (int i = 0; i < NP; i++) pn[i] = msk[i] ? p[i] + x[i] : p[i];
        const double rms_n = sqrt(rms_at(pn));
        bool done = false;
        if (rms_n < rms) {
            good = true;
            done = rms_n >= rms * (1.0 - A.rms_diff);
            rms = rms_n;
#pragma unroll
            for (int i = 0; i < NP; i++) p[i] = pn[i];
            if (!done && iter + 1 < A.num_iter) normal_at(p);   // the next step solves from here
        } else {
            good = false;
        }
        if (done) break;
        if (good) {
            lambda *= A.l_good;
        } else {
            lambda *= A.l_bad;
            if (lambda > A.l_max) break;
        }
    }
    // runLma (Correlations2dLMA:510-524): success when the last step improved or the final rms is
    // below the initial one (an improving step always lowers it)
    out.ok = good || (rms < rms0);
    if (out.ok) {   // getDisparityStrength (Correlations2dLMA:165-185)
        double sum_amp = 0.0;
        if (ex0) sum_amp += st->gwn[0] * p[NC];
        if (ex1) sum_amp += st->gwn[1] * (ex0 ? p[NC + 1] : p[NC]);
        const double max_amp = fmax(vmax, 0.0);
        if (sum_amp > 1.25 * max_amp) sum_amp = max_amp;
        out.disp = -p[0];
        out.strength = sum_amp;
        out.wm = p[1];
        out.wxy = p[P_WXY];
    }
    return out;
}

// C2:1663-1750 foregroundCorrect; returns the disparity to use
__device__ double lma_foreground(const LmaArgs& A, const bool* have, const double (*dds)[4], double full_disp) {
    const bool bg = A.fo_far != 0, ortho = A.fo_ortho != 0;
    double mn = NAN, mx = NAN, eff[4], width_d[4];
    int imn = -1, imx = -1;
    bool strong[4];
    for (int i = 0; i < 4; i++) {
        eff[i] = NAN;
        width_d[i] = 0.0;
        strong[i] = false;
        if (!have[i]) continue;
        if (isnan(mn) || dds[i][0] < mn) { mn = dds[i][0]; imn = i; }
        if (isnan(mx) || dds[i][0] > mx) { mx = dds[i][0]; imx = i; }
        eff[i] = dds[i][1] * (dds[i][2] - dds[i][3]) / (dds[i][2] + dds[i][3]);
        width_d[i] = dds[i][2] + dds[i][3];
        strong[i] = (dds[i][1] >= A.fo_min_strength) && (eff[i] >= A.fo_min_eff);
    }
    const bool are_ortho = (imn ^ imx) == 1;
    double corr = A.fo_overcorr * (mx - mn);
    const int isel = bg ? imn : imx, iother = bg ? imx : imn;
    if (isel < 0 || !strong[isel]) corr = NAN;
    else if (width_d[isel] > A.fo_max_hwidth) corr = NAN;
    else if (eff[isel] / eff[isel ^ 1] < A.fo_min_eff_ratio) corr = NAN;
    else {
        if (ortho && !are_ortho) corr = 0.0;
        else if (iother < 0 || !strong[iother]) corr = 0.0;
        if (!(mx >= (mn + A.fo_min_diff))) corr = NAN;
    }
    double disp = full_disp;
    if (!isnan(corr)) {
        if (bg) disp = fmax(mn - corr, full_disp - (full_disp - mn) * A.fo_lim_overcorr);
        else disp = fmin(mx + corr, full_disp + (mx - full_disp) * A.fo_lim_overcorr);
    }
    return disp;
}

// ID:2254-2336 over the task list, as a chain of five kernels on one stream: the samples of each
// tile's fit over all pairs (k_lma_setup_all), its iterations (k_lma_iterate), the samples of the
// four one-direction fits (k_lma_setup_dirs), their iterations, and the foreground correction
// (k_lma_finish).  Every fit's state goes through device memory (LmaState, 328 bytes), so the
// set-up kernels keep the tile's lanes busy on the per-cell work at their own occupancy and the
// iterations run one thread per fit, without a wavefront of a workgroup waiting for another.
// The records are read at CLT_DISP_CM_RAW / CLT_DISP_STR_RAW and written at CM, STRENGTH (final),
// POLY, POLY_STR.  NC = 4 when the extra vertical width Wyd is adjusted, else 3.  dir_out (the
// rig's tileCorrs, ID:7581-7740): when not null, the four one-direction fits run on every tile
// whose fit over all pairs succeeded and [T][8] receives their disparities (NaN when a fit fails or
// does not run) and strengths (0 then); no foreground correction.

// k_lma_setup_all: 16 lanes per tile, 4 tiles per wavefront; state of tile t at st[t].
template <int NC>
__global__ __launch_bounds__(256, 4) void k_lma_setup_all(const float* __restrict__ corr, const float* __restrict__ disp,
                                                       int num_tasks, const LmaArgs* __restrict__ Ap,
                                                       LmaState* __restrict__ st) {
    const LmaArgs& A = *Ap;   // in device memory: the fits index it per lane
    __shared__ double s_v[LMA_SLOTS][2 * LMA_CELLS];
    __shared__ uint16_t s_idx[LMA_SLOTS][2 * LMA_CELLS];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int sub = lane >> 4, l = lane & 15, slot = wave * LMA_TILES_PER_WAVE + sub;
    const int t = (xcd_block(blockIdx.x, gridDim.x) * LMA_WAVES + wave) * LMA_TILES_PER_WAVE + sub;
    if (t >= num_tasks) return;   // whole rows: the reductions never see an exited lane
    const float* rec = disp + (size_t)t * CLT_DISP_LAYERS;
    const float cm = rec[CLT_DISP_CM_RAW], str = rec[CLT_DISP_STR_RAW];
    // the wavefront's first tile as a 64-bit base, the tile within it as a small 32-bit offset
    const float* corr_w = corr + (size_t)__builtin_amdgcn_readfirstlane(t - sub) * A.npairs * CLT_CORR_LEN;
    const int toff = sub * A.npairs * CLT_CORR_LEN;
    if (!isnan(cm) && (double)str > A.min_poly_strength) {
        const LmaLds S = {s_v[slot], s_idx[slot], &st[t]};
        lma_fit<16, NC, LMA_GPAIRS>(A, &A.fit[0], corr_w, toff, -(double)cm, l, lane, 0xffffull << (16 * sub), S);
    } else if (l == 0) {
        st[t].run = 0;
    }
}

// k_lma_iterate: one thread per fit; res[i] = {ok, disparity, strength, Wm, Wxy}.
template <int NC>
__global__ __launch_bounds__(256) void k_lma_iterate(const LmaState* __restrict__ st, int num_fits,
                                                     const LmaArgs* __restrict__ Ap, double* __restrict__ res) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_fits) return;
    LmaOut r = {false, 0.0, 0.0, 0.0, 0.0};
    if (st[i].run) r = lma_iterate<NC>(*Ap, &st[i]);
    double* o = res + (size_t)i * 5;
    o[0] = r.ok ? 1.0 : 0.0;
    o[1] = r.disp;
    o[2] = r.strength;
    o[3] = r.wm;
    o[4] = r.wxy;
}

// k_lma_setup_dirs: the result of the fit over all pairs (res_all[t]) mixed into the tile's values
// (ID:2273-2281); when the foreground correction can apply, the samples of the four one-direction
// fits (corr4dirsLMA, C2:1591-1630), two at a time on 8 lanes each, states at st[4 t + direction];
// tmp[t] = {disparity, strength, CM out, strength out, POLY, POLY strength, directions run}.
template <int NC>
__global__ __launch_bounds__(256, 3) void k_lma_setup_dirs(const float* __restrict__ corr, const float* __restrict__ disp,
                                                        int num_tasks, const LmaArgs* __restrict__ Ap,
                                                        const double* __restrict__ res_all, LmaState* __restrict__ st,
                                                        double* __restrict__ tmp, int all_dirs) {
    const LmaArgs& A = *Ap;
    __shared__ double s_v[LMA_SLOTS][2 * LMA_CELLS];
    __shared__ uint16_t s_idx[LMA_SLOTS][2 * LMA_CELLS];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int sub = lane >> 4, l = lane & 15, slot = wave * LMA_TILES_PER_WAVE + sub;
    const int t = (xcd_block(blockIdx.x, gridDim.x) * LMA_WAVES + wave) * LMA_TILES_PER_WAVE + sub;
    if (t >= num_tasks) return;   // whole rows
    const float* rec = disp + (size_t)t * CLT_DISP_LAYERS;
    const float cm = rec[CLT_DISP_CM_RAW], str = rec[CLT_DISP_STR_RAW];
    const float* corr_w = corr + (size_t)__builtin_amdgcn_readfirstlane(t - sub) * A.npairs * CLT_CORR_LEN;
    const int toff = sub * A.npairs * CLT_CORR_LEN;
    double out_cm = cm, out_st = str, poly = NAN, poly_s = NAN;
    double disparity = cm, strength = str;
    bool dirs = false;
    const double* r = res_all + (size_t)t * 5;
    if (!isnan(cm) && (double)str > A.min_poly_strength && r[0] != 0.0) {
        poly = r[1];
        poly_s = r[2];
        if (A.mix) {
            disparity = r[1];
            strength = r[2];
            out_cm = disparity;
            out_st = strength;
        }
        // the foreground correction is kept only when strength > fo_min_strength (ID:2319-2322):
        // below that the four fits cannot change the result
        dirs = all_dirs ? true : (A.fo_correct && strength > fmax(0.0, A.fo_min_strength));
    }
    if (dirs) {
        const int half = l >> 3, hl = l & 7;
        for (int pass = 0; pass < 2; pass++) {
            const int f = 1 + 2 * pass + half;
            const LmaLds S = {s_v[slot] + half * LMA_CELLS, s_idx[slot] + half * LMA_CELLS, &st[4 * (size_t)t + f - 1]};
            lma_fit<8, NC, LMA_DPAIRS>(A, &A.fit[f], corr_w, toff, -disparity, hl, lane,
                                       0xffull << (16 * sub + 8 * half), S);
            wave_sync();
        }
    } else if (l < 4) {
        st[4 * (size_t)t + l].run = 0;
    }
    if (l == 0) {
        double* o = tmp + (size_t)t * 8;
        o[0] = disparity;
        o[1] = strength;
        o[2] = out_cm;
        o[3] = out_st;
        o[4] = poly;
        o[5] = poly_s;
        o[6] = dirs ? 1.0 : 0.0;
    }
}

// k_lma_finish: one thread per tile: the foreground correction (ID:2288-2336) from the four fits'
// results res_dirs[4 t + direction] (or their disparities and strengths to dir_out), then the record.
__global__ __launch_bounds__(256) void k_lma_finish(float* __restrict__ disp, int num_tasks,
                                                    const LmaArgs* __restrict__ Ap, const double* __restrict__ tmp,
                                                    const double* __restrict__ res_dirs, float* __restrict__ dir_out) {
    const LmaArgs& A = *Ap;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_tasks) return;
    const double* o = tmp + (size_t)t * 8;
    const double disparity = o[0], strength = o[1];
    double out_cm = o[2];
    if (o[6] != 0.0) {
        bool have[4];
        double dv[4][4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const double* r = res_dirs + (4 * (size_t)t + d) * 5;
            dv[d][0] = (r[0] != 0.0) ? r[1] : NAN;
            dv[d][1] = r[2];
            dv[d][2] = r[3];
            dv[d][3] = r[4];
            have[d] = !isnan(dv[d][0]);
        }
        if (dir_out) {
#pragma unroll
            for (int d = 0; d < 4; d++) {
                dir_out[(size_t)t * 8 + d] = (float)dv[d][0];
                dir_out[(size_t)t * 8 + 4 + d] = have[d] ? (float)dv[d][1] : 0.f;
            }
        } else {
            const double mod = lma_foreground(A, have, dv, disparity);
            if (mod != disparity && strength > A.fo_min_strength) out_cm = mod;
        }
    } else if (dir_out) {
#pragma unroll
        for (int d = 0; d < 4; d++) {
            dir_out[(size_t)t * 8 + d] = NAN;
            dir_out[(size_t)t * 8 + 4 + d] = 0.f;
        }
    }
    float* rec = disp + (size_t)t * CLT_DISP_LAYERS;
    rec[CLT_DISP_CM] = (float)out_cm;
    rec[CLT_DISP_STRENGTH] = (float)o[3];
    rec[CLT_DISP_POLY] = (float)o[4];
    rec[CLT_DISP_POLY_STR] = (float)o[5];
}

// tileInterCamCorrs (ID:7370-7560) after the centre of mass: the LMA of the inter-camera
// correlation as one horizontal pair (C2:1903-1935) and its dx, dy.  rec_in [T][8] from
// k_corr_inter: {disparity, strength, LMA start (NaN: no centre of mass), -, CM dx, CM dy, -, -};
// res_out [T][4] = {disparity, strength, dx, dy}.  16 lanes per tile as k_corr_lma.
template <int NC>
__global__ __launch_bounds__(256, 3) void k_corr_lma_inter(const float* __restrict__ corr, const float* __restrict__ rec_in,
                                                        float* __restrict__ res_out, int num_tasks,
                                                        const LmaArgs* __restrict__ Ap, int use_poly, int use_xy_poly,
                                                        int notch) {
    const LmaArgs& A = *Ap;
    __shared__ double s_v[LMA_SLOTS][LMA_CELLS];
    __shared__ uint16_t s_idx[LMA_SLOTS][LMA_CELLS];
    __shared__ LmaState s_st[LMA_SLOTS];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int sub = lane >> 4, l = lane & 15, slot = wave * LMA_TILES_PER_WAVE + sub;
    const int t = (xcd_block(blockIdx.x, gridDim.x) * LMA_WAVES + wave) * LMA_TILES_PER_WAVE + sub;
    if (t >= num_tasks) return;   // whole rows: the reductions never see an exited lane
    const float* r = rec_in + (size_t)t * 8;
    double disparity = r[0], strength = r[1], dx = NAN, dy = NAN;
    const float x0 = r[2];
    if (!isnan(x0) && strength > A.min_poly_strength && !notch) {
        const LmaLds S = {s_v[slot], s_idx[slot], &s_st[slot]};
        const uint64_t fmask = 0xffffull << (16 * sub);
        const float* corr_w = corr + (size_t)__builtin_amdgcn_readfirstlane(t - sub) * CLT_CORR_LEN;   // wave base
        lma_fit<16, NC, 1>(A, &A.fit[LMA_FIT_SINGLE], corr_w, sub * CLT_CORR_LEN, (double)x0, l, lane, fmask, S);
        wave_sync();
        LmaOut q = {false, 0.0, 0.0, 0.0, 0.0};
        if (s_st[slot].run) q = lma_iterate<NC, false>(A, &s_st[slot]);   // the same fit on every lane of the tile
        if (q.ok) {
            if (use_poly) {
                disparity = q.disp;
                strength = q.strength;
            }
            if (use_xy_poly) {   // single2dPoly (C2:1509-1536) around 0
                const LmaOut pq = lma_fit<16, NC, 1, true>(A, &A.fit[LMA_FIT_SINGLE], corr_w, sub * CLT_CORR_LEN, 0.0,
                                                           l, lane, fmask, S);
                if (pq.ok) {
                    dx = -pq.disp;
                    dy = -pq.wm;
                }
            } else {   // single2dCM (C2:1538-1586), computed with the centre of mass
                dx = r[4];
                dy = r[5];
            }
        }
    }
    if (l == 0) {
        float* o = res_out + (size_t)t * 4;
        o[0] = (float)disparity;
        o[1] = (float)strength;
        o[2] = (float)dx;
        o[3] = (float)dy;
    }
}

}  // namespace clt
