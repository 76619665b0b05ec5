// Per-tile correlation LMA and foreground correction (gfx950).
//
// The reference refines every tile's disparity after the centre-of-mass step (ImageDtt.java
// ID:2254-2330): a Levenberg-Marquardt fit of a paraboloid to the combined pair correlations
// (Correlation2d.corrLMA, C2:1754-1900; sample selection filterConvex C2:2029-2207 and addSamples
// C2:2230-2338; the fit Correlations2dLMA.java), four one-direction fits (corr4dirsLMA,
// C2:1591-1630) and the foreground correction (foregroundCorrect, C2:1663-1750).
//
// Mapping: 16 lanes (one DPP row) per tile, 4 tiles per wavefront.  A fit's samples are the cells
// of the 9x9 box around the start point: each lane owns box cells l, l+16, ... (6 per group), the
// combined correlation values and sample weights of the tile sit in LDS, the selection grows in
// LDS byte flags, and every normal-equation sum is an all-reduce over the row (DPP row rotates on
// the two halves of each double; every lane ends with bit-identical sums, so the fit's decisions
// are uniform within the tile).  Arithmetic is FP64 like the reference; the 6x6 normal system is
// solved by LDL^T in registers (each lane redundantly).
#pragma once

namespace clt {

constexpr int LMA_TILES_PER_WAVE = 4;
constexpr int LMA_WAVES = 4;                 // 256-thread blocks
constexpr int LMA_SLOTS = LMA_TILES_PER_WAVE * LMA_WAVES;
constexpr int LMA_CELLS = 81;                // 9x9 box: half window at most 4
constexpr int LMA_CPL = 6;                   // box cells per lane (16 * 6 >= 81)
constexpr int LMA_MAXG = 2;                  // compatible pair groups per fit
constexpr int LMA_GPAIRS = 8;                // pairs per group
constexpr int LMA_NACC = 28;                 // 21 (upper triangle of 6x6) + 6 + rms
constexpr int LMA_NFITS = 5;                 // all pairs, then horizontal, vertical, diag main, diag other

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

// ---- row-of-16 reductions on doubles (DPP row rotates; identical result in all 16 lanes) ----------
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double row_sum(double v) {
    v += dpp_f64<0x128>(v);   // row_ror:8
    v += dpp_f64<0x124>(v);   // row_ror:4
    v += dpp_f64<0x122>(v);   // row_ror:2
    v += dpp_f64<0x121>(v);   // row_ror:1
    return v;
}
__device__ __forceinline__ double row_max(double v) {
    v = fmax(v, dpp_f64<0x128>(v));
    v = fmax(v, dpp_f64<0x124>(v));
    v = fmax(v, dpp_f64<0x122>(v));
    v = fmax(v, dpp_f64<0x121>(v));
    return v;
}
// any lane of this tile's row
__device__ __forceinline__ bool row_any(bool p, int sub) {
    return ((__ballot(p) >> (16 * sub)) & 0xffffull) != 0;
}

// index of (i, j), i <= j, in the packed upper triangle of the 6x6 normal matrix
__host__ __device__ constexpr int lma_tri(int i, int j) { return i * 6 - i * (i - 1) / 2 + (j - i); }

struct LmaOut {
    bool ok;
    double disp, strength, wm, wxy;
};

// One Correlations2dLMA fit of tile t (its 16 lanes) over the groups of F, started at xcenter
// (pixels along the disparity axis, = -disparity).  C2:1754-1900.
__device__ __forceinline__ LmaOut lma_fit(const LmaArgs& A, const LmaFit& F, const float* __restrict__ tcorr,
                                       double xcenter, int l, int sub, double2* s_vw, uint8_t* s_new,
                                       uint8_t* s_cvx, double* s_acc) {
    LmaOut out = {false, 0.0, 0.0, 0.0, 0.0};
    const int G = F.ngroups;
    if (G == 0) return out;
    const int hwin = A.hwin;
    unsigned smp_bits = 0;                       // [group][k] own cells that are samples
    double gw_raw[LMA_MAXG] = {0.0, 0.0};        // per group: sum of raw sample weights (group_weights)
    double sw_fit = 0.0, vmax = -INFINITY;
    int x0s[LMA_MAXG], y0s[LMA_MAXG];
#pragma unroll
    for (int g = 0; g < LMA_MAXG; g++) {
        if (g >= G) break;
        const LmaGroup& grp = F.g[g];
        const int ixc = (int)floor(xcenter / grp.scale + 0.5);   // Math.round (C2:2256)
        const int x0 = ixc + 7, y0 = grp.diag ? ixc + 7 : 7;
        x0s[g] = x0;
        y0s[g] = y0;
        double2* vw = s_vw + g * LMA_CELLS;
        // combined correlation of the group's pairs over the box (combineCompatiblePairs, C2:462-488)
#pragma unroll 1
        for (int k = 0; k < LMA_CPL; k++) {
            const int c = l + 16 * k;
            if (c >= LMA_CELLS) break;
            const int row = y0 - 4 + c / 9, col = x0 - 4 + c % 9;
            double v = 0.0;
            if (row >= 0 && row < 15 && col >= 0 && col < 15) {
                for (int p = 0; p < grp.npairs; p++) {
                    const int xf = grp.xform[p];
                    const int src = (xf == 1) ? col * 15 + row : ((xf == 2) ? (14 - row) * 15 + col : row * 15 + col);
                    v += (double)tcorr[grp.pair[p] * CLT_CORR_LEN + src];
                }
                if (grp.npairs > 1) v /= grp.npairs;
            }
            vw[c].x = v;
        }
        wave_sync();
        // bi-convex cells inside the clamped window (C2:2046-2064)
        const int min_row = max(y0 - hwin, 0), max_row = min(y0 + hwin, 14);
        const int min_col = max(x0 - hwin, 0), max_col = min(x0 + hwin, 14);
        unsigned own_cvx = 0, sel = 0;
        const bool start_in = (y0 >= min_row && y0 <= max_row && x0 >= min_col && x0 <= max_col);
#pragma unroll 1
        for (int k = 0; k < LMA_CPL; k++) {
            const int c = l + 16 * k;
            if (c >= LMA_CELLS) break;
            const int row = y0 - 4 + c / 9, col = x0 - 4 + c % 9;
            bool cv = false;
            if (row > min_row && row < max_row && col > min_col && col < max_col) {
                const double d = vw[c].x;
                cv = (2 * d > (vw[c - 1].x + vw[c + 1].x)) && (2 * d > (vw[c - 9].x + vw[c + 9].x));
            }
            own_cvx |= (unsigned)cv << k;
            s_cvx[c] = cv;
            const bool st = start_in && c == 40;
            sel |= (unsigned)st << k;
            s_new[c] = st;
        }
        wave_sync();
        // grow the selection ring by ring (C2:2073-2195); sel / s_new hold the same set between steps
        for (int istep = 0; istep < hwin; istep++) {
            const bool first = istep == 0;
            const int smin_row = max(y0 - istep, 0), smax_row = min(y0 + istep, 14);
            const int smin_col = max(x0 - istep, 0), smax_col = min(x0 + istep, 14);
#pragma unroll 1
            for (int k = 0; k < LMA_CPL; k++) {
                const int c = l + 16 * k;
                if (c >= LMA_CELLS) break;
                if (!((sel >> k) & 1) || !(((own_cvx >> k) & 1) || first)) continue;
                const int row = y0 - 4 + c / 9, col = x0 - 4 + c % 9;
                if (istep < y0 && row == y0 - istep && col >= smin_col && col <= smax_col) {        // up
                    const int j = c - 9;
                    s_new[j] = 1;
                    if (s_cvx[j]) {
                        if (col <= x0 && col > min_col) s_new[j - 1] = 1;
                        if (col >= x0 && col < max_col) s_new[j + 1] = 1;
                    }
                }
                if (y0 + istep < max_row && row == y0 + istep && col >= smin_col && col <= smax_col) {   // down
                    const int j = c + 9;
                    s_new[j] = 1;
                    if (s_cvx[j]) {
                        if (col <= x0 && col > min_col) s_new[j - 1] = 1;
                        if (col >= x0 && col < max_col) s_new[j + 1] = 1;
                    }
                }
                if (istep < x0 && col == x0 - istep && row >= smin_row && row <= smax_row) {        // left
                    const int j = c - 1;
                    s_new[j] = 1;
                    if (s_cvx[j]) {
                        if (row <= y0 && row > min_row) s_new[j - 9] = 1;
                        if (row >= y0 && row < max_row) s_new[j + 9] = 1;
                    }
                }
                if (x0 + istep < max_col && col == x0 + istep && row >= smin_row && row <= smax_row) {   // right
                    const int j = c + 1;
                    s_new[j] = 1;
                    if (s_cvx[j]) {
                        if (row <= y0 && row > min_row) s_new[j - 9] = 1;
                        if (row >= y0 && row < max_row) s_new[j + 9] = 1;
                    }
                }
            }
            wave_sync();
            unsigned nsel = 0;
#pragma unroll 1
            for (int k = 0; k < LMA_CPL; k++) {
                const int c = l + 16 * k;
                if (c >= LMA_CELLS) break;
                nsel |= (unsigned)(s_new[c] != 0) << k;
            }
            const bool grew = row_any((nsel & ~sel) != 0, sub);
            sel = nsel;
            if (!grew) break;
        }
        wave_sync();
        // cell weights (C2:2197-2206) and samples (addSamples, C2:2230-2338)
        double gsum = 0.0;
#pragma unroll 1
        for (int k = 0; k < LMA_CPL; k++) {
            const int c = l + 16 * k;
            if (c >= LMA_CELLS) break;
            const int br = c / 9, bc = c % 9;
            const int row = y0 - 4 + br, col = x0 - 4 + bc;
            const bool ing = row >= 0 && row < 15 && col >= 0 && col < 15;
            double w = ((sel >> k) & 1) ? (((own_cvx >> k) & 1) ? 1.0 : A.nc_cost) : 0.0;
            if (A.add3x3 && ing && br >= 3 && br <= 5 && bc >= 3 && bc <= 5 && w == 0.0) w = A.nc_cost;
            w *= grp.npairs;
            double wv = 0.0, vv = 0.0;
            if (ing && w > 0.0) {
                smp_bits |= 1u << (g * LMA_CPL + k);
                const double v = vw[c].x;
                if (A.vasw_pwr != 0.0) w *= (A.vasw_pwr == 2.0) ? v * v : pow(fabs(v), A.vasw_pwr);
                gsum += w;
                if (!isnan(v) && !isnan(w)) {
                    wv = w;
                    vv = v;
                }
                sw_fit += wv;
                vmax = fmax(vmax, vv);
            }
            vw[c] = make_double2(vv, wv);
        }
        gw_raw[g] = row_sum(gsum);
    }
    // groups are numbered in the order of their first sample (Correlations2dLMA:196-210)
    bool gex[LMA_MAXG];
    int ag_of[LMA_MAXG], ng = 0;
#pragma unroll
    for (int g = 0; g < LMA_MAXG; g++) {
        gex[g] = g < G && row_any(((smp_bits >> (g * LMA_CPL)) & 0x3fu) != 0, sub);
        ag_of[g] = gex[g] ? 4 + ng++ : -1;
    }
    if (ng == 0) return out;
    sw_fit = row_sum(sw_fit);
    vmax = row_max(vmax);
    // setWeightsValues (Correlations2dLMA:245-278): weights normalised to a unit sum with the two
    // regularisation points (Wm -> half_width at cost_wm, Wxy -> 0 at cost_wxy)
    const double tot = sw_fit + A.cost_wm + A.cost_wxy;
    const double inv = (tot != 0.0) ? 1.0 / tot : 1.0;
    const double wr0 = A.cost_wm * inv, wr1 = A.cost_wxy * inv;
    double gwn[LMA_MAXG];
#pragma unroll
    for (int g = 0; g < LMA_MAXG; g++) gwn[g] = (sw_fit > 0.0) ? gw_raw[g] / sw_fit : gw_raw[g];
#pragma unroll
    for (int g = 0; g < LMA_MAXG; g++)
#pragma unroll 1
        for (int k = 0; k < LMA_CPL; k++) {
            if (g >= G) break;
            const int c = l + 16 * k;
            if (c >= LMA_CELLS) break;
            s_vw[g * LMA_CELLS + c].y *= inv;
        }
    wave_sync();
    // parameters: X0, Wm, Wyd, Wxy, Ag0, Ag1 (initVector, Correlations2dLMA:212-243)
    bool msk[6] = {true, A.adj_wm != 0, A.adj_wyd != 0, A.adj_wxy != 0, ng > 0 && A.adj_ag != 0,
                   ng > 1 && A.adj_ag != 0};
    double p[6] = {xcenter, A.half_width, 0.0, 0.0, vmax, vmax};
    double acc[LMA_NACC];
    const double sqrt2 = 1.4142135623730951;

    // one pass over the samples at parameters q: normal matrix, J^T W (y - f) and the weighted
    // residual sum (getFxJt + getWJtJlambda + getWYmFxRms, Correlations2dLMA:368-461)
    auto evaluate = [&](const double* q) {
#pragma unroll
        for (int i = 0; i < LMA_NACC; i++) acc[i] = 0.0;
#pragma unroll
        for (int g = 0; g < LMA_MAXG; g++) {
            if (!gex[g]) continue;
            const LmaGroup& grp = F.g[g];
            const double ws = grp.diag ? sqrt2 : 1.0;
            const double Wy = q[1] * grp.scale + q[2], Wx = Wy + q[3];
            const double iWx = 1.0 / Wx, iWy = 1.0 / Wy;
            const double Ag = (ag_of[g] == 4) ? q[4] : q[5];
            const bool a0 = ag_of[g] == 4;
            const int x0 = x0s[g], y0 = y0s[g];
#pragma unroll 1
            for (int k = 0; k < LMA_CPL; k++) {
                const int c = l + 16 * k;
                if (c >= LMA_CELLS) break;
                if (!((smp_bits >> (g * LMA_CPL + k)) & 1)) continue;
                const double2 s = s_vw[g * LMA_CELLS + c];
                const int cx = x0 - 11 + c % 9, cy = y0 - 11 + c / 9;   // relative to the correlation centre
                const double x = grp.diag ? 0.5 * (cx + cy) * grp.scale : cx * grp.scale;
                const double y = grp.diag ? 0.5 * (cy - cx) * grp.scale : cy * grp.scale;
                const double dxw = ws * (x - q[0]) * iWx, dyw = ws * y * iWy;
                const double d = 1.0 - dxw * dxw - dyw * dyw;
                const double f = d * Ag;
                const double dfdWx = Ag * 2 * dxw * dxw * iWx, dfdWy = Ag * 2 * dyw * dyw * iWy;
                double j[6];
                j[0] = ws * Ag * 2 * dxw * iWx;
                j[1] = grp.scale * (dfdWx + dfdWy);
                j[2] = dfdWx + dfdWy;
                j[3] = dfdWx;
                j[4] = a0 ? d : 0.0;
                j[5] = a0 ? 0.0 : d;
                const double r = s.x - f, wr = s.y * r;
                acc[27] += wr * r;
#pragma unroll
                for (int a = 0; a < 6; a++) {
                    acc[21 + a] += j[a] * wr;
                    const double wj = s.y * j[a];
#pragma unroll
                    for (int b = a; b < 6; b++) acc[lma_tri(a, b)] += wj * j[b];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < LMA_NACC; i++) acc[i] = row_sum(acc[i]);
        // the two regularisation points
        const double r0 = A.half_width - q[1], r1 = -q[3];
        acc[lma_tri(1, 1)] += wr0;
        acc[lma_tri(3, 3)] += wr1;
        acc[22] += wr0 * r0;
        acc[24] += wr1 * r1;
        acc[27] += wr0 * r0 * r0 + wr1 * r1 * r1;
    };

    // Iteration -1 evaluates the start point; every later one solves the damped normal equations
    // of the accepted point (its sums in LDS, ping-pong buffers cur / 1 - cur) and evaluates the
    // trial point (lmaStep, Correlations2dLMA:533-640; runLma 477-530).
    double pn[6];
#pragma unroll
    for (int i = 0; i < 6; i++) pn[i] = p[i];
    double rms0 = 0.0, rms = 0.0, lambda = A.lambda0;
    bool good = false;
    int cur = 0;
    for (int iter = -1; iter < A.num_iter; iter++) {
        if (iter >= 0) {
            // (J^T W J + lambda diag(J^T W J)) delta = J^T W (y - f) over the adjusted parameters, LDL^T
            const double* sa = s_acc + cur * 32;
            double L[6][6], dd[6], x[6];
            bool singular = false;
#pragma unroll
            for (int k = 0; k < 6; k++) {
                const double v = sa[lma_tri(k, k)];
                double dk = msk[k] ? v + v * lambda : 1.0;
#pragma unroll
                for (int s2 = 0; s2 < k; s2++) dk -= L[k][s2] * L[k][s2] * dd[s2];
                dd[k] = dk;
                singular |= (dk == 0.0);
                const double idk = 1.0 / dk;
#pragma unroll
                for (int i = k + 1; i < 6; i++) {
                    double w = (msk[i] && msk[k]) ? sa[lma_tri(k, i)] : 0.0;
#pragma unroll
                    for (int s2 = 0; s2 < k; s2++) w -= L[i][s2] * L[k][s2] * dd[s2];
                    L[i][k] = w * idk;
                }
            }
            if (singular) break;   // lmaStep returns {false, true}
#pragma unroll
            for (int i = 0; i < 6; i++) {
                double w = msk[i] ? sa[21 + i] : 0.0;
#pragma unroll
                for (int s2 = 0; s2 < i; s2++) w -= L[i][s2] * x[s2];
                x[i] = w;
            }
#pragma unroll
            for (int i = 5; i >= 0; i--) {
                double w = x[i] / dd[i];
#pragma unroll
                for (int s2 = i + 1; s2 < 6; s2++) w -= L[s2][i] * x[s2];
                x[i] = w;
            }
#pragma unroll
            for (int i = 0; i < 6; i++) pn[i] = msk[i] ? p[i] + x[i] : p[i];
        }
        evaluate(pn);
        const double rms_n = sqrt(acc[27]);
        const int trial = iter < 0 ? 0 : 1 - cur;
        if (l == 0)
#pragma unroll
            for (int i = 0; i < LMA_NACC; i++) s_acc[trial * 32 + i] = acc[i];
        wave_sync();
        if (iter < 0) {
            rms0 = rms = rms_n;
            cur = trial;
            continue;
        }
        bool done = false;
        if (rms_n < rms) {
            good = true;
            done = rms_n >= rms * (1.0 - A.rms_diff);
            rms = rms_n;
            cur = trial;
#pragma unroll
            for (int i = 0; i < 6; i++) p[i] = pn[i];
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
    // below the initial one; an improving step always lowers it, so the second test covers both
    out.ok = good || (rms < rms0);
    if (out.ok) {   // getDisparityStrength (Correlations2dLMA:165-185)
        double sum_amp = 0.0;
#pragma unroll
        for (int g = 0; g < LMA_MAXG; g++)
            if (gex[g]) sum_amp += gwn[g] * ((ag_of[g] == 4) ? p[4] : p[5]);
        const double max_amp = fmax(vmax, 0.0);
        if (sum_amp > 1.25 * max_amp) sum_amp = max_amp;
        out.disp = -p[0];
        out.strength = sum_amp;
        out.wm = p[1];
        out.wxy = p[3];
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

// ID:2254-2330 over the task list: reads the raw centre of mass and strength of each record,
// writes CM, STRENGTH (final), POLY, POLY_STR.  16 lanes per tile, 4 tiles per wavefront.
__global__ __launch_bounds__(256, 2) void k_corr_lma(const float* __restrict__ corr, float* __restrict__ disp,
                                                  int num_tasks, LmaArgs A) {
    __shared__ double2 s_vw[LMA_SLOTS][LMA_MAXG * LMA_CELLS];
    __shared__ uint8_t s_new[LMA_SLOTS][LMA_CELLS + 15];
    __shared__ uint8_t s_cvx[LMA_SLOTS][LMA_CELLS + 15];
    __shared__ double s_acc[LMA_SLOTS][64];
    __shared__ double s_dds[LMA_SLOTS][4][4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int sub = lane >> 4, l = lane & 15, slot = wave * LMA_TILES_PER_WAVE + sub;
    const int t = (xcd_block(blockIdx.x, gridDim.x) * LMA_WAVES + wave) * LMA_TILES_PER_WAVE + sub;
    if (t >= num_tasks) return;   // whole rows: the row reductions never see an exited lane
    float* rec = disp + (size_t)t * CLT_DISP_LAYERS;
    const float cm = rec[CLT_DISP_CM_RAW], st = rec[CLT_DISP_STR_RAW];
    const float* tcorr = corr + (size_t)t * A.npairs * CLT_CORR_LEN;
    double out_cm = cm, out_st = st, poly = NAN, poly_s = NAN;
    double disparity = cm, strength = st, xc = -(double)cm;
    bool run = !isnan(cm) && (double)st > A.min_poly_strength, dirs = false;
    double (*dds)[4] = s_dds[slot];
    // fit 0 over all oriented pairs, then (when it succeeded) the four one-direction fits
    for (int j = 0; j < LMA_NFITS; j++) {
        LmaOut r = {false, 0.0, 0.0, 0.0, 0.0};
        if (run) r = lma_fit(A, A.fit[j], tcorr, xc, l, sub, s_vw[slot], s_new[slot], s_cvx[slot], s_acc[slot]);
        if (j == 0) {
            if (!r.ok) break;
            poly = r.disp;
            poly_s = r.strength;
            if (A.mix) {
                disparity = r.disp;
                strength = r.strength;
                out_cm = disparity;
                out_st = strength;
            }
            run = dirs = A.fo_correct && strength > 0.0;
            xc = -disparity;
        } else if (l == 0) {
            dds[j - 1][0] = r.ok ? r.disp : NAN;
            dds[j - 1][1] = r.strength;
            dds[j - 1][2] = r.wm;
            dds[j - 1][3] = r.wxy;
        }
    }
    if (dirs) {   // foregroundCorrect (ID:2288-2330)
        wave_sync();
        bool have[4];
        double dv[4][4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
#pragma unroll
            for (int e = 0; e < 4; e++) dv[d][e] = dds[d][e];
            have[d] = !isnan(dv[d][0]);
        }
        const double mod = lma_foreground(A, have, dv, disparity);
        if (mod != disparity && strength > A.fo_min_strength) out_cm = mod;
    }
    if (l == 0) {
        rec[CLT_DISP_CM] = (float)out_cm;
        rec[CLT_DISP_STRENGTH] = (float)out_st;
        rec[CLT_DISP_POLY] = (float)poly;
        rec[CLT_DISP_POLY_STR] = (float)poly_s;
    }
}

}  // namespace clt
