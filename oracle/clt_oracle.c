/*
 * clt_oracle.c -- CPU restatement of imagej-elphel's CLT tile processor (the reference CPU path).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle of the MI355X tile processor: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it, and only as the checker
 * (or as the timed CPU baseline).  The product library never links or calls it.
 *
 * It restates, function by function, the Java CPU path of the reference (paths relative to
 * src/main/java/com/elphel/imagej/tileprocessor/):
 *   DttRad2.java       (DR) transforms, window, fold/unfold tables, correlation unfold
 *   ImageDtt.java      (ID) extract_correct_tile, convolve_tile, fract_shift, clt_lpf, iclt_2d,
 *                           the quad correlation/texture driver, tile_debayer, tile_combine_rgba,
 *                           combineRGBATiles, the worker-thread pool
 *   Correlation2d.java (C2) phase correlation, strips, strip combination, maxima
 * Inputs are the float32 buffers of the C ABI (include/clt_tile_processor.h): mosaic images, kernels
 * in the GPU upload layout, packed task records.  Arithmetic is done in `real`: double by default
 * (the FP64 oracle), float when compiled with -DORC_FP32 (an FP32 port used to size the expected
 * single-precision error).  Elementary functions are always evaluated in double.
 *
 * Compile with -ffp-contract=off: Java never fuses multiply-add, and the integer tile positions
 * (ID:4940-4943) are recomputed bit-exactly by the GPU from the same float inputs.
 *
 * Parity pinning of every function is listed in the header of tests/test_oracle.py.
 */
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/clt_tile_processor.h"

#ifdef ORC_FP32
typedef float real;
#else
typedef double real;
#endif
/* Port weights of the texture combiner are double in both builds: in float, exp(-k d) underflows
 * to 0 for every port of a pixel where all ports disagree, and the reference's double arithmetic
 * (ID:2880-2935) keeps their ratios there. */
typedef double wreal;

#define EXPORT __attribute__((visibility("default")))
#define N8 8

/* ================================ DttRad2 tables ================================ */
static double HW[8];                       /* DR:671-694 set_window(1): sin(pi*(i+0.5)/16)      */
static double CIV[8][8], SIV[8][8];        /* DR:993-1012, DR:1054-1073 (orthonormal, symmetric)  */
static double CIIE[8][8], SIIE[8][8];      /* DR:952-968 (setup_CIIe), DR:1014-1030 (setup_SIIe)  */
static double CIIIE[8][8], SIIIE[8][8];    /* DR:970-991 (setup_CIIIe), DR:1032-1052 (setup_SIIIe)*/
static int    FOLD_INDEX[64][4];           /* DR:146-203 set_fold_2d                             */
static int    UNFOLD_INDEX[256];           /* DR:405-434 set_unfold_2d                           */
static double UNFOLD_K[4][256];
static pthread_once_t tables_once = PTHREAD_ONCE_INIT;

/* DR:112-144 get_fold_indices(x, n=8) */
static void get_fold_indices(int x, int ind[2][6]) {
    const int n = 8, n1 = 4;
    if (x < n1) {
        ind[0][0] = n + n1 - x - 1; ind[0][1] = n1 + x;     ind[0][2] = -1; ind[0][3] = 1;
        ind[0][4] = n1 - x - 1;     ind[0][5] = -1;
        ind[1][0] = n + n1 + x;     ind[1][1] = n1 - x - 1; ind[1][2] = -1; ind[1][3] = -1;
        ind[1][4] = n1 + x;         ind[1][5] = -1;
    } else {
        x -= n1;
        ind[0][0] = x;              ind[0][1] = x;          ind[0][2] = 1;  ind[0][3] = 1;
        ind[0][4] = n - x - 1;      ind[0][5] = 1;
        ind[1][0] = n - x - 1;      ind[1][1] = n - x - 1;  ind[1][2] = -1; ind[1][3] = 1;
        ind[1][4] = x;              ind[1][5] = 1;
    }
}

/* DR:392-403 get_unfold_index_signs(x, n=8) */
static void get_unfold_index_signs(int x, int is2[3]) {
    const int n = 8, n1 = 4;
    int segm = x / n1;
    x = x % n1;
    switch (segm) {
    case 0: is2[0] = x + n1;     is2[1] = 1;  is2[2] = 1;  break;
    case 1: is2[0] = n - x - 1;  is2[1] = -1; is2[2] = 1;  break;
    case 2: is2[0] = n1 - x - 1; is2[1] = -1; is2[2] = 1;  break;
    default: is2[0] = x;         is2[1] = -1; is2[2] = -1; break;
    }
}

static void init_tables(void) {
    const int n = 8;
    for (int i = 0; i < n; i++) HW[i] = sin(M_PI / (2.0 * n) * (i + 0.5));
    const double scale = sqrt(2.0 / n);
    for (int j = 0; j < n; j++) {
        for (int k = 0; k < n; k++) {
            CIV[j][k] = scale * cos((2 * j + 1) * (2 * k + 1) * M_PI / (4 * n));
            SIV[j][k] = scale * sin((2 * j + 1) * (2 * k + 1) * M_PI / (4 * n));
            CIIE[j][k] = scale * cos(j * (2 * k + 1) * M_PI / (2 * n));
            SIIE[j][k] = scale * sin((j + 1) * (2 * k + 1) * M_PI / (2 * n));
        }
    }
    for (int j = 0; j < n; j++) {           /* CIIIe[k][j], e_j = 0.5 for j == 0 or n-1 (DR:982) */
        double ec = ((j == 0) || (j == n - 1)) ? 0.5 : 1.0;
        double es = (j == n - 1) ? 0.5 : 1.0;  /* SIIIe: only j == n-1 (DR:1043) */
        for (int k = 0; k < n; k++) {
            CIIIE[k][j] = scale * ec * cos(j * (2 * k + 1) * M_PI / (2 * n));
            SIIIE[k][j] = scale * es * sin((j + 1) * (2 * k + 1) * M_PI / (2 * n));
        }
    }
    /* DR:146-203 set_fold_2d: fold_index[n*i + j][k] = 2n*vert_ind[k>>1] + hor_ind[k&1] */
    for (int i = 0; i < n; i++) {
        int fv[2][6];
        get_fold_indices(i, fv);
        for (int j = 0; j < n; j++) {
            int fh[2][6];
            get_fold_indices(j, fh);
            for (int k = 0; k < 4; k++)
                FOLD_INDEX[n * i + j][k] = 2 * n * fv[(k >> 1) & 1][0] + fh[k & 1][0];
        }
    }
    /* DR:405-434 set_unfold_2d */
    for (int i = 0; i < 2 * n; i++) {
        int iv[3];
        get_unfold_index_signs(i, iv);
        double wv = HW[(i < n) ? i : 2 * n - i - 1];
        double kv[2] = {iv[1] * wv, iv[2] * wv};
        for (int j = 0; j < 2 * n; j++) {
            int ih[3];
            get_unfold_index_signs(j, ih);
            double wh = HW[(j < n) ? j : 2 * n - j - 1];
            double kh[2] = {ih[1] * wh, ih[2] * wh};
            UNFOLD_INDEX[2 * n * i + j] = iv[0] * n + ih[0];
            for (int mode = 0; mode < 4; mode++)
                UNFOLD_K[mode][2 * n * i + j] = kv[(mode >> 1) & 1] * kh[mode & 1];
        }
    }
}
static void tables(void) { pthread_once(&tables_once, init_tables); }

/* ================================ 8-point transforms ================================ */
/* 1-D y = M x for one of the tables (dct_iv/dst_iv DR:80-98 equal the orthonormal closed forms) */
static void mat8(const double m[8][8], const real* x, real* y) {
    for (int i = 0; i < 8; i++) {
        real s = 0;
        for (int j = 0; j < 8; j++) s += (real)m[i][j] * x[j];
        y[i] = s;
    }
}

/* Separable 2-D pass pair with the reference's transposed storage (DR:503-519 dttt_iv, DR:596-614
 * dttt_iie, DR:648-666 dttt_iiie): first pass along the contiguous index with table a0/a1 by
 * mode bit 0, result transposed, second pass with b by mode bit 1. */
static void dttt_2d(const real* x, int mode, const double (*c)[8], const double (*s)[8], real* y) {
    real t[64], line[8];
    for (int i = 0; i < 8; i++) {
        mat8((mode & 1) ? s : c, x + 8 * i, line);
        for (int j = 0; j < 8; j++) t[j * 8 + i] = line[j];
    }
    for (int i = 0; i < 8; i++) mat8((mode & 2) ? s : c, t + 8 * i, y + 8 * i);
}
static void dttt_iv(const real* x, int mode, real* y) { dttt_2d(x, mode, CIV, SIV, y); }
static void dttt_iie(const real* x, int mode, real* y) { dttt_2d(x, mode, CIIE, SIIE, y); }
static void dttt_iiie(const real* x, int mode, real* y) { dttt_2d(x, mode, CIIIE, SIIIE, y); }

/* DR:263-314 get_shifted_fold_2d(8, shift_hor, shift_vert) */
static void shifted_fold_coeffs(real shift_hor, real shift_vert, real fk[4][64][4]) {
    const double f = M_PI / 16.0;
    real ahc = (real)cos(f * shift_hor), ahs = (real)sin(f * shift_hor);
    real avc = (real)cos(f * shift_vert), avs = (real)sin(f * shift_vert);
    for (int i = 0; i < 8; i++) {
        int fi[2][6];
        get_fold_indices(i, fi);
        real vw0 = avc * (real)HW[fi[0][1]] + avs * (real)HW[fi[0][4]] * fi[0][5];
        real vw1 = avc * (real)HW[fi[1][1]] + avs * (real)HW[fi[1][4]] * fi[1][5];
        real vk[2][2] = {{fi[0][2] * vw0, fi[1][2] * vw1}, {fi[0][3] * vw0, fi[1][3] * vw1}};
        for (int j = 0; j < 8; j++) {
            int fj[2][6];
            get_fold_indices(j, fj);
            real hw0 = ahc * (real)HW[fj[0][1]] + ahs * (real)HW[fj[0][4]] * fj[0][5];
            real hw1 = ahc * (real)HW[fj[1][1]] + ahs * (real)HW[fj[1][4]] * fj[1][5];
            real hk[2][2] = {{fj[0][2] * hw0, fj[1][2] * hw1}, {fj[0][3] * hw0, fj[1][3] * hw1}};
            for (int mode = 0; mode < 4; mode++)
                for (int k = 0; k < 4; k++)
                    fk[mode][8 * i + j][k] = vk[(mode >> 1) & 1][(k >> 1) & 1] * hk[mode & 1][k & 1];
        }
    }
}

/* DR:724-738 fold_tile: 16x16 [row][col] -> 8x8 [v][h] */
static void fold_tile(const real* x, int mode, real fk[4][64][4], real* y) {
    for (int i = 0; i < 64; i++) {
        real s = 0;
        for (int k = 0; k < 4; k++) s += x[FOLD_INDEX[i][k]] * fk[mode][i][k];
        y[i] = s;
    }
}

/* DR:770-780 unfold_tile: 8x8 [v][h] IDCT output -> 16x16 [row][col] */
static void unfold_tile(const real* x, int mode, real* y) {
    for (int i = 0; i < 256; i++) y[i] = (real)UNFOLD_K[mode][i] * x[UNFOLD_INDEX[i]];
}

/* DR:443-489 corr_unfold_tile: 4 quadrants [y][x] of DCT-IIe/DST-IIe output -> 15x15 */
static void corr_unfold_tile(real q[4][64], real* r) {
    const int n = 8, cs = 15;
    const real sc = (real)0.25;
    r[cs * n - n] = sc * q[0][0];
    for (int j = 1; j < n; j++) {
        r[cs * n - n + j] = sc * (q[0][j] + q[1][j - 1]);
        r[cs * n - n - j] = sc * (q[0][j] - q[1][j - 1]);
    }
    for (int i = 1; i < n; i++) {
        r[cs * (n + i) - n] = sc * (q[0][i * n] + q[2][(i - 1) * n]);
        r[cs * (n - i) - n] = sc * (q[0][i * n] - q[2][(i - 1) * n]);
        for (int j = 1; j < n; j++) {
            real a = q[0][i * n + j], b = q[1][i * n + j - 1], c = q[2][(i - 1) * n + j],
                 d = q[3][(i - 1) * n + j - 1];
            r[cs * (n + i) - n + j] = sc * (a + b + c + d);
            r[cs * (n + i) - n - j] = sc * (a + -b + c + -d);
            r[cs * (n - i) - n + j] = sc * (a + b + -c + -d);
            r[cs * (n - i) - n - j] = sc * (a + -b + -c + d);
        }
    }
}

/* ================================ LPF ================================ */
/* ID:1662-1690 (filter in clt_aberrations_quad_corr_new) == ID:4178-4220 (clt_lpf) == GTP:814-848
 * (setCltLpf): exp(-(i^2+j^2)/(2*sigma)) (sigma not squared, as in the reference), normalised by
 * the DC gain of the DCT-IIIe, then dttt_iiie and * 2n. */
static void make_lpf(double sigma, real* out) {
    real fd[64], f[64];
    if (sigma == 0) {
        for (int i = 0; i < 64; i++) fd[i] = (i == 0) ? 1 : 0;
    } else {
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 8; j++) fd[i * 8 + j] = (real)exp(-(i * i + j * j) / (2 * sigma));
    }
    real sum = 0;
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            real d = fd[i * 8 + j];
            d *= (real)(cos(M_PI * i / 16.0) * cos(M_PI * j / 16.0));
            if (i > 0) d *= 2;
            if (j > 0) d *= 2;
            sum += d;
        }
    for (int i = 0; i < 64; i++) fd[i] /= sum;
    dttt_iiie(fd, 0, f);
    for (int i = 0; i < 64; i++) out[i] = f[i] * 16;
}

/* ================================ convert (ID:4867-5584) ================================ */
typedef struct {
    const clt_geom* g;
    int tiles_x, tiles_y, num_tiles;
} frame_t;

static void frame_init(frame_t* f, const clt_geom* g) {
    f->g = g;
    f->tiles_x = g->width / 8;
    f->tiles_y = g->height / 8;
    f->num_tiles = f->tiles_x * f->tiles_y;
}

/* Colour of mosaic pixel (EyesisCorrections.java:2219 for the reference phase red_row=0,
 * red_col=1; blue sits at the opposite parities, green on the remaining diagonal). */
static int pixel_color(const clt_geom* g, int y, int x) {
    if (g->num_colors == 1) return 0;
    int ry = (y & 1) == g->red_row, rx = (x & 1) == g->red_col;
    if (ry && rx) return 0;
    if (!ry && !rx) return 1;
    return 2;
}

/* Java Math.round(double) */
static double jround(double x) { return floor(x + 0.5); }

/* ID:4867-5426 extract_correct_tile followed by fract_shift (ID:2047-2058) for one camera/colour.
 * image: mosaic [height][stride]; kernels/offsets: GPU upload layout or NULL.  clt: [4][64] in the
 * reference CPU order [h][v]; returns residual shift in res[2]. */
static void fract_shift(real clt[4][64], real shift_x, real shift_y);
static void convolve_tile(real data[4][64], const real kern[4][64]);

static void convert_one(const clt_geom* g, const float* image, int stride, const float* kernels,
                        const float* offsets, int color, double center_x, double center_y,
                        real clt[4][64], double res[2]) {
    const int width = g->width, height = g->height;
    double px = center_x - 8, py = center_y - 8;
    int use_kernels = (kernels != NULL) && (offsets != NULL);
    const float* ktile = NULL;
    if (use_kernels) {
        const int kstep = g->kernel_step;
        int ktx = (int)jround(center_x / kstep) + 1;
        int kty = (int)jround(center_y / kstep) + 1;
        if (kty < 0) kty = 0; else if (kty >= g->kernels_vert) kty = g->kernels_vert - 1;
        if (ktx < 0) ktx = 0; else if (ktx >= g->kernels_hor) ktx = g->kernels_hor - 1;
        size_t kidx = ((size_t)kty * g->kernels_hor + ktx) * g->num_colors + color;
        const float* ce = offsets + kidx * 8;   /* CltExtra: data_x,data_y,center_x,center_y,
                                                    dxc_dx,dxc_dy,dyc_dx,dyc_dy (TPC:139-148) */
        double kdx = center_x - (ktx - 1 + 0.5) * kstep;
        double kdy = center_y - (kty - 1 + 0.5) * kstep;
        px = center_x - 8 - ((double)ce[0] + (double)ce[4] * kdx + (double)ce[5] * kdy);
        py = center_y - 8 - ((double)ce[1] + (double)ce[6] * kdx + (double)ce[7] * kdy);
        ktile = kernels + kidx * 256;
    }
    int left = (int)-jround(-px), top = (int)-jround(-py);  /* ID:4940-4941 */
    res[0] = -(px - left);
    res[1] = -(py - top);
    real tile_in[256];
    for (int i = 0; i < 16; i++) {                    /* ID:4953-4970, per-pixel parity clamp */
        int pi = top + i;
        if (pi < 0) pi &= 1;
        else if (pi >= height) pi = height - 2 + (pi & 1);
        for (int j = 0; j < 16; j++) {
            int pj = left + j;
            if (pj < 0) pj &= 1;
            else if (pj >= width) pj = width - 2 + (pj & 1);
            tile_in[16 * i + j] = (pixel_color(g, pi, pj) == color) ? (real)image[(size_t)pi * stride + pj] : 0;
        }
    }
    real fk[4][64][4];
    shifted_fold_coeffs((real)res[0], (real)res[1], fk);  /* ID:5071 */
    for (int m = 0; m < 4; m++) {                         /* ID:5243-5292 */
        real folded[64];
        fold_tile(tile_in, m, fk, folded);
        dttt_iv(folded, m, clt[m]);
    }
    if (use_kernels) {                                    /* ID:5402-5410: [v][h] upload -> [h][v] */
        real kern[4][64];
        for (int m = 0; m < 4; m++)
            for (int v = 0; v < 8; v++)
                for (int h = 0; h < 8; h++) kern[m][h * 8 + v] = (real)ktile[m * 64 + v * 8 + h];
        convolve_tile(clt, kern);
    }
}

/* ID:5429-5488 convolve_tile */
static void convolve_tile(real data[4][64], const real kern[4][64]) {
    static const int zi[4][4] = {{0, -1, -2, 3}, {1, 0, -3, -2}, {2, -3, 0, -1}, {3, 2, 1, 0}};
    real r[4][64];
    for (int i = 0; i < 64; i++)
        for (int n = 0; n < 4; n++) {
            real s = 0;
            for (int k = 0; k < 4; k++) {
                if (zi[n][k] < 0) s -= data[-zi[n][k]][i] * kern[k][i];
                else s += data[zi[n][k]][i] * kern[k][i];
            }
            r[n][i] = s;
        }
    memcpy(data, r, sizeof(r));
}

/* ID:5490-5584 fract_shift: phase rotation, horizontal then vertical ([h][v] storage) */
static void fract_shift(real clt[4][64], real shift_x, real shift_y) {
    real ch[8], sh[8], cv[8], sv[8];
    for (int i = 0; i < 8; i++) {
        ch[i] = (real)cos((i + 0.5) * M_PI * shift_x / 8);
        sh[i] = (real)sin((i + 0.5) * M_PI * shift_x / 8);
        cv[i] = (real)cos((i + 0.5) * M_PI * shift_y / 8);
        sv[i] = (real)sin((i + 0.5) * M_PI * shift_y / 8);
    }
    for (int idx = 0; idx < 64; idx++) {
        int kh = idx >> 3, kv = idx & 7;
        real c = ch[kh], s = sh[kh];
        real t0 = clt[0][idx] * c - clt[1][idx] * s;
        real t1 = clt[1][idx] * c + clt[0][idx] * s;
        real t2 = clt[2][idx] * c - clt[3][idx] * s;
        real t3 = clt[3][idx] * c + clt[2][idx] * s;
        c = cv[kv]; s = sv[kv];
        clt[0][idx] = t0 * c - t2 * s;
        clt[2][idx] = t2 * c + t0 * s;
        clt[1][idx] = t1 * c - t3 * s;
        clt[3][idx] = t3 * c + t1 * s;
    }
}

/* ================================ thread pool (ID:5847-5872) ================================ */
typedef void (*tile_fn)(void* ctx, int index);
typedef struct {
    atomic_int next;
    int count;
    tile_fn fn;
    void* ctx;
} pool_t;

static void* pool_worker(void* arg) {
    pool_t* p = (pool_t*)arg;
    for (int i = atomic_fetch_add(&p->next, 1); i < p->count; i = atomic_fetch_add(&p->next, 1))
        p->fn(p->ctx, i);
    return NULL;
}

static void run_pool(int count, int nthreads, tile_fn fn, void* ctx) {
    pool_t p;
    atomic_init(&p.next, 0);
    p.count = count;
    p.fn = fn;
    p.ctx = ctx;
    if (nthreads <= 1 || count < 2) {
        pool_worker(&p);
        return;
    }
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    int started = 0;
    for (int t = 0; t < nthreads; t++) {
        if (pthread_create(&th[t], NULL, pool_worker, &p) != 0) break;
        started++;
    }
    if (started == 0) pool_worker(&p);
    for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
}

/* ================================ task records (GTP:132-166) ================================ */
static int task_words(const clt_geom* g) { return 2 + 2 * g->num_cams; }
static int32_t task_flags(const clt_geom* g, const int32_t* tasks, int t) { return tasks[(size_t)t * task_words(g)]; }
static void task_txy(const clt_geom* g, const int32_t* tasks, int t, int* tx, int* ty) {
    int32_t txy = tasks[(size_t)t * task_words(g) + 1];
    *tx = txy & 0xffff;
    *ty = (txy >> 16) & 0xffff;
}
static float task_xy(const clt_geom* g, const int32_t* tasks, int t, int cam, int d) {
    float v;
    memcpy(&v, &tasks[(size_t)t * task_words(g) + 2 + 2 * cam + d], 4);
    return v;
}

/* ================================ frame: convert ================================ */
typedef struct {
    frame_t f;
    const float* const* images;
    int stride;
    const float* const* kernels;
    const float* const* offsets;
    const int32_t* tasks;
    int lpf_mask;
    const double* lpf;     /* [3][64] or NULL */
    int no_fract_shift;
    double* clt;           /* [cam][tile][color][4][v][h] */
    double* residual;      /* [task][cam][color][2] or NULL */
} convert_ctx;

static void convert_task(void* vctx, int t) {
    convert_ctx* c = (convert_ctx*)vctx;
    const clt_geom* g = c->f.g;
    if (task_flags(g, c->tasks, t) == 0) return;
    int tx, ty;
    task_txy(g, c->tasks, t, &tx, &ty);
    size_t tile = (size_t)ty * c->f.tiles_x + tx;
    for (int cam = 0; cam < g->num_cams; cam++) {
        double cx = task_xy(g, c->tasks, t, cam, 0), cy = task_xy(g, c->tasks, t, cam, 1);
        for (int col = 0; col < g->num_colors; col++) {
            real clt[4][64];
            double res[2];
            convert_one(g, c->images[cam], c->stride, c->kernels ? c->kernels[cam] : NULL,
                        c->offsets ? c->offsets[cam] : NULL, col, cx, cy, clt, res);
            if (!c->no_fract_shift) fract_shift(clt, (real)res[0], (real)res[1]);
            if (c->lpf && ((c->lpf_mask >> col) & 1))
                for (int m = 0; m < 4; m++)
                    for (int i = 0; i < 64; i++) clt[m][i] *= (real)c->lpf[col * 64 + i];
            double* o = c->clt + (((size_t)cam * c->f.num_tiles + tile) * g->num_colors + col) * 256;
            for (int m = 0; m < 4; m++)
                for (int v = 0; v < 8; v++)
                    for (int h = 0; h < 8; h++) o[m * 64 + v * 8 + h] = clt[m][h * 8 + v];
            if (c->residual) {
                double* r = c->residual + (((size_t)t * g->num_cams + cam) * g->num_colors + col) * 2;
                r[0] = res[0];
                r[1] = res[1];
            }
        }
    }
}

/* Read one camera/colour CLT tile of the GPU layout back into [4][h*8+v] */
static void load_clt(const frame_t* f, const double* clt, int cam, size_t tile, int col, real out[4][64]) {
    const double* p = clt + (((size_t)cam * f->num_tiles + tile) * f->g->num_colors + col) * 256;
    for (int m = 0; m < 4; m++)
        for (int v = 0; v < 8; v++)
            for (int h = 0; h < 8; h++) out[m][h * 8 + v] = (real)p[m * 64 + v * 8 + h];
}

/* ================================ correlation (C2) ================================ */
static const int ZI[4][4] = {{0, 1, 2, 3}, {-1, 0, -3, 2}, {-2, -3, 0, 1}, {3, -2, -1, 0}}; /* C2:86-90 */

/* C2:197-231 correlateSingleColorFD */
static void correlate_single_color_fd(real d1[4][64], real d2[4][64], real fat_zero, real tcorr[4][64]) {
    real a2[64], sa2 = 0;
    for (int i = 0; i < 64; i++) {
        real s1 = 0, s2 = 0;
        for (int n = 0; n < 4; n++) {
            s1 += d1[n][i] * d1[n][i];
            s2 += d2[n][i] * d2[n][i];
        }
        a2[i] = (real)sqrt((double)s1 * (double)s2);   /* in double as C2:213 (no float underflow) */
        sa2 += a2[i];
    }
    real fz2 = sa2 / 64 * fat_zero * fat_zero;
    for (int i = 0; i < 64; i++) {
        real scale = 1 / (a2[i] + fz2);
        for (int n = 0; n < 4; n++) {
            real s = 0;
            for (int k = 0; k < 4; k++) {
                if (ZI[n][k] < 0) s -= d1[-ZI[n][k]][i] * d2[k][i];
                else s += d1[ZI[n][k]][i] * d2[k][i];
            }
            tcorr[n][i] = s * scale;
        }
    }
}

/* C2:275-324 correlateCompositeFD (one pair): colour mix, LPF, DCT-IIe/DST-IIe, unfold 15x15 */
static void correlate_composite_fd(int ncol, real (*clt1)[4][64], real (*clt2)[4][64], const real* lpf,
                                   const real* col_weights, real fat_zero, real* out225) {
    real acc[4][64], t[4][64];
    for (int col = 0; col < ncol; col++) {
        correlate_single_color_fd(clt1[col], clt2[col], fat_zero, t);
        real w = (ncol == 1) ? 1 : col_weights[col];
        for (int n = 0; n < 4; n++)
            for (int i = 0; i < 64; i++) {
                if (col == 0) acc[n][i] = t[n][i] * w;
                else acc[n][i] += t[n][i] * w;
            }
    }
    if (lpf)
        for (int n = 0; n < 4; n++)
            for (int i = 0; i < 64; i++) acc[n][i] *= lpf[i];
    real q[4][64];
    for (int quadrant = 0; quadrant < 4; quadrant++) {
        int mode = ((quadrant << 1) & 2) | ((quadrant >> 1) & 1);
        dttt_iie(acc[quadrant], mode, q[quadrant]);
    }
    corr_unfold_tile(q, out225);
}

/* C2:610-718 scaleRotateInterpoateSingleCorrelation */
static void strip_single(const real* corr, int hwidth, int dir, int ss, real* strip) {
    const int center = 7, width = 15;
    int denom = ss * ((dir > 1) ? 1 : 2);
    real rdenom = (real)denom;
    int ilimit = center * denom;
    for (int i = 0; i < hwidth * width; i++) strip[i] = 0;   /* Java arrays start zeroed */
    for (int row = 0; row < hwidth; row++) {
        for (int scol = -center; scol <= center; scol++) {
            int indx = row * width + scol + center;
            for (int down = -1; down < ((row == 0) ? 0 : 2); down += 2) {
                int xnum = 0, ynum = 0;
                switch (dir) {
                case 0: xnum = 2 * scol + (row & 1); ynum = down * row; break;
                case 1: xnum = down * row; ynum = 2 * scol + (row & 1); break;
                case 2: xnum = scol - ((down * row) >> 1); ynum = scol + ((down * row + 1) >> 1); break;
                case 3: xnum = scol - ((down * row) >> 1); ynum = -scol - ((down * row + 1) >> 1); break;
                }
                if ((xnum > ilimit) || (xnum < -ilimit) || (ynum > ilimit) || (ynum < -ilimit)) {
                    strip[indx] = (real)NAN;
                } else {
                    int ix0 = (xnum + ilimit) / denom;
                    int iy0 = (ynum + ilimit) / denom;
                    int int_x = (xnum % denom) == 0;
                    int int_y = (ynum % denom) == 0;
                    real d;
                    if (int_x && int_y) {
                        d = corr[iy0 * width + ix0];
                    } else {
                        real dx = (xnum + ilimit) / rdenom - ix0;
                        real dy = (ynum + ilimit) / rdenom - iy0;
                        if (int_y) d = (1 - dx) * corr[iy0 * width + ix0] + dx * corr[iy0 * width + ix0 + 1];
                        else if (int_x) d = (1 - dy) * corr[iy0 * width + ix0] + dy * corr[iy0 * width + ix0 + width];
                        else
                            d = ((1 - dx) * (1 - dy) * corr[iy0 * width + ix0]) + (dx * (1 - dy) * corr[iy0 * width + ix0 + 1]) +
                                ((1 - dx) * dy * corr[iy0 * width + ix0 + width]) + (dx * dy * corr[iy0 * width + ix0 + width + 1]);
                    }
                    if (row == 0) strip[indx] = d;
                    else strip[indx] += (real)0.5 * d;
                }
            }
        }
    }
}

/* C2:730-790 combineInterpolatedCorrelations */
static int combine_strips(real** strips, int nstrips, int len, real offset, int twice_diagonal, real* combo) {
    int ncombined = 0, have = 0;
    if (offset >= 0) {
        for (int p = 0; p < nstrips; p++) {
            if (!strips[p]) continue;
            if (!have) {
                for (int i = 0; i < len; i++) combo[i] = 1;
                have = 1;
            }
            if (twice_diagonal) {
                for (int i = 0; i < len; i++) {
                    real d = strips[p][i] + offset;
                    if (d < 0) d = 0;
                    combo[i] *= d * d;
                }
                ncombined++;
            } else {
                for (int i = 0; i < len; i++) {
                    real d = strips[p][i] + offset;
                    if (d < 0) d = 0;
                    combo[i] *= d;
                }
            }
            ncombined++;
        }
        double pwr = 1.0 / ncombined;
        if (have)
            for (int i = 0; i < len; i++)
                if (combo[i] > 0) combo[i] = (real)pow((double)combo[i], pwr) - offset;
    } else {
        for (int p = 0; p < nstrips; p++) {
            if (!strips[p]) continue;
            if (!have) {
                for (int i = 0; i < len; i++) combo[i] = 0;
                have = 1;
            }
            if (twice_diagonal) {
                for (int i = 0; i < len; i++) combo[i] += 2 * strips[p][i];
                ncombined++;
            } else {
                for (int i = 0; i < len; i++) combo[i] += strips[p][i];
            }
            ncombined++;
        }
        real scale = (real)1 / ncombined;
        for (int i = 0; i < len; i++) combo[i] *= scale;
    }
    return have;
}

/* C2:802-843 getMaxXYInt, axis only (rectangular strip).  Returns 1 and ix (relative to centre). */
static int max_xy_int_axis(const real* data, real min_max, int* ix) {
    int imx = 0;
    for (int i = 1; i < 15; i++)
        if (isnan(data[imx]) || (data[i] > data[imx])) imx = i;
    if (!(data[imx] >= min_max)) return 0;
    *ix = imx % 15 - 7;
    return 1;
}

/* C2:1047-1081 halfFlatTopWindow */
static void half_flat_top_window(int ihwidth, double hwidth, double blur, int normalize, int notch,
                                 double scale, double* wnd) {
    for (int i = 0; i < ihwidth; i++) {
        if (blur <= 0) wnd[i] = (i < hwidth) ? 1.0 : 0.0;
        else if (i < hwidth - blur / 2) wnd[i] = 1.0;
        else if (i > hwidth + blur / 2) wnd[i] = 0.0;
        else wnd[i] = 0.5 * (1.0 - sin(M_PI * (i - hwidth) / blur));
        if (notch) wnd[i] = 1.0 - wnd[i];
    }
    if (normalize) {
        double s = 0;
        for (int i = 0; i < ihwidth; i++) s += ((i == 0) ? 1.0 : 2.0) * wnd[i];
        s = 1 / s;
        for (int i = 0; i < ihwidth; i++) wnd[i] *= s;
    }
    for (int i = 0; i < ihwidth; i++) wnd[i] *= scale;
}

/* C2:936-1032 getMaxXCm on a strip with rows, window_y/window_x from IDP defaults (C2:139-152) */
static int max_x_cm(const real* data, int data_height, int ixcenter, const double* wy_tab, int wy_len,
                    const double* wx_tab, int wx_len, real rslt[3]) {
    const int center = 7, data_width = 15;
    real wy_scale = 1;
    if (data_height > wy_len) {
        data_height = wy_len;
    } else if (data_height < wy_len) {
        real swy = (real)wy_tab[0];
        for (int i = 1; i < data_height; i++) swy += (real)wy_tab[i];
        wy_scale = 1 / swy;
    }
    real s0 = 0, sx = 0, sx2 = 0;
    int x0 = center + ixcenter;
    for (int dy = 0; dy < data_height; dy++) {
        int odd = dy & 1;
        real wy = ((dy == 0) ? wy_scale : (2 * wy_scale)) * (real)wy_tab[dy];
        int indx0 = data_width * dy;
        for (int adx = odd; adx < wx_len; adx += 2) {
            for (int dir = (adx == 0) ? 1 : -1; dir <= 1; dir += 2) {
                int idx = (adx * dir) >> 1;
                int x = 2 * idx + odd;
                int x1 = x0 + idx;
                if ((x1 >= 0) && (x1 < data_width)) {
                    real d = data[indx0 + x1];
                    if (!isnan(d) && (d > 0)) {
                        d *= wy * (real)wx_tab[adx];
                        s0 += d;
                        sx += d * x;
                        sx2 += d * x * x;
                    }
                }
            }
        }
    }
    if (s0 == 0) return 0;
    rslt[0] = ixcenter + sx / s0 / 2;
    rslt[1] = s0;
    rslt[2] = (real)sqrt((double)(s0 * sx2 - sx * sx)) / s0 / 2;
    return 1;
}

/* C2:2654-2675 createOrtoNotch */
static void ortho_notch(double enhortho_width, double enhortho_scale, double* f) {
    for (int i = 0; i < 15; i++) {
        if ((i < (8 - enhortho_width)) || (i > (8 - 2 + enhortho_width))) f[i] = 1.0;
        else f[i] = enhortho_scale;
        if (i == 7) f[i] = 0.0;
        f[i] *= sin(M_PI * (i + 1.0) / 16);
    }
}

/* C2:2502-2651 getMaxXSOrtho (offset >= 0 branch and linear branch), 3-point parabola
 * (PolynomialApproximation.polynomialApproximation1d with 3 points, degree 2, closed form). */
static int max_xs_ortho(real** corrs, int npairs, real offset, int is_vert, const double* notch, real rslt[2]) {
    const int cs = 15;
    if (npairs == 0) return 0;
    real corr_1d[15];
    double w = 1.0 / npairs;
    for (int j = 0; j < cs; j++) {
        corr_1d[j] = 0;
        if (offset >= 0) {
            for (int i = 0; i < cs; i++) {
                real combo = 1;
                for (int p = 0; p < npairs; p++) {
                    const real* c = corrs[p];
                    real d = is_vert ? (real)0.5 * (c[j * cs + i] + c[j * cs + (cs - i - 1)])
                                     : (real)0.5 * (c[i * cs + j] + c[(cs - i - 1) * cs + j]);
                    combo *= d + offset;
                }
                if (combo > 0) combo = (real)pow((double)combo, w) - offset;
                else combo = -offset;
                corr_1d[j] += combo * (real)notch[i];
            }
        } else {
            for (int p = 0; p < npairs; p++)
                for (int i = 0; i < cs; i++) {
                    const real* c = corrs[p];
                    corr_1d[j] += (real)w * (is_vert ? c[j * cs + i] : c[i * cs + j]) * (real)notch[i];
                }
        }
    }
    int icenter = 0;
    for (int i = 1; i < cs; i++)
        if (corr_1d[i] > corr_1d[icenter]) icenter = i;
    real xcenter = (real)icenter;
    if ((icenter > 0) && (icenter < cs - 1)) {
        real a = corr_1d[icenter - 1], b = corr_1d[icenter], c = corr_1d[icenter + 1];
        real den = a - 2 * b + c;
        if (den != 0) xcenter = icenter + (a - c) / (2 * den);
    }
    icenter = (int)jround((double)xcenter);
    rslt[0] = xcenter - 7;
    rslt[1] = corr_1d[icenter] / 8;
    return 1;
}

/* ============ per-tile correlation LMA (Correlation2d.java, Correlations2dLMA.java) ============ */
/* The fit always runs in double (the reference's arithmetic), also in the -DORC_FP32 build. */
#define LMA_HWIN_MAX 4
#define LMA_MAX_GROUPS 6
#define LMA_BOX ((2 * LMA_HWIN_MAX + 1) * (2 * LMA_HWIN_MAX + 1))
#define LMA_MAX_SAMPLES (LMA_MAX_GROUPS * LMA_BOX)
#define LMA_MAX_PARS (4 + LMA_MAX_GROUPS)
#define LMA_MAX_POINTS (LMA_MAX_SAMPLES + 2)
enum { LMA_X0 = 0, LMA_WM = 1, LMA_WYD = 2, LMA_WXY = 3, LMA_AG = 4 };   /* Correlations2dLMA:61-65 */

/* C2:72-78 GROUPS: {diagonal, baseline scale} in the reference order */
static const int LMA_GROUPS[6][2] = {{0, 1}, {1, 1}, {0, 2}, {1, 2}, {0, 4}, {1, 4}};

static int pair_diag(const clt_pair* p) { return p->dir == 2 || p->dir == 3; }
static int pair_compatible(const clt_pair* p, int diag, int scale) {
    return p->dir >= 0 && pair_diag(p) == diag && p->scale == scale;
}

/* C2:462-488 combineCompatiblePairs: horizontal and main-diagonal pairs as they are, vertical pairs
 * transposed (getTransposeAllOrtho, C2:168-177), other-diagonal pairs flipped upside down
 * (getTransposeAllDiagonal, C2:179-188); the sum is divided by the number of pairs. */
static int combine_compatible(const double (*corrs)[225], const clt_pair* pairs, int npairs, const char* sel,
                              int diag, int scale, double* combo) {
    int n = 0;
    for (int i = 0; i < 225; i++) combo[i] = 0;
    for (int p = 0; p < npairs; p++) {
        if (!sel[p] || !pair_compatible(&pairs[p], diag, scale)) continue;
        for (int r = 0; r < 15; r++)
            for (int c = 0; c < 15; c++) {
                int src = r * 15 + c;
                if (pairs[p].dir == 1) src = c * 15 + r;
                else if (pairs[p].dir == 3) src = (14 - r) * 15 + c;
                combo[r * 15 + c] += corrs[p][src];
            }
        n++;
    }
    if (n > 1)
        for (int i = 0; i < 225; i++) combo[i] /= n;
    return n;
}

/* C2:2029-2207 filterConvex: grow the selection from (x0c, y0c) ring by ring through bi-convex
 * cells; selected cells weigh 1 (convex) or nc_cost, the 3x3 around the start at least nc_cost. */
static void filter_convex(const double* d, int hwin, int x0c, int y0c, int add3x3, double nc_cost, double* weights) {
    const int center = 7, width = 15;
    const int x0 = x0c + center, y0 = y0c + center;
    char convex[225] = {0}, sel[225] = {0}, sel_new[225];
    const int min_row = (y0 - hwin < 0) ? 0 : y0 - hwin, max_row = (y0 + hwin >= width) ? width - 1 : y0 + hwin;
    const int min_col = (x0 - hwin < 0) ? 0 : x0 - hwin, max_col = (x0 + hwin >= width) ? width - 1 : x0 + hwin;
    for (int row = min_row + 1; row < max_row; row++)
        for (int col = min_col + 1; col < max_col; col++) {
            int i = row * width + col;
            convex[i] = (2 * d[i] > (d[i - 1] + d[i + 1])) && (2 * d[i] > (d[i - width] + d[i + width]));
        }
    if (y0 >= min_row && y0 <= max_row && x0 >= min_col && x0 <= max_col) sel[y0 * width + x0] = 1;
#define IN225(i) ((i) >= 0 && (i) < 225)
#define MARK(i) do { if (!sel_new[i]) { sel_new[i] = 1; num_new++; } } while (0)
    for (int istep = 0; istep < hwin; istep++) {
        memcpy(sel_new, sel, sizeof(sel));
        const int first = (istep == 0);
        int num_new = 0;
        const int smin_row = (y0 - istep < 0) ? 0 : y0 - istep, smax_row = (y0 + istep >= width) ? width - 1 : y0 + istep;
        const int smin_col = (x0 - istep < 0) ? 0 : x0 - istep, smax_col = (x0 + istep >= width) ? width - 1 : x0 + istep;
        if (istep < y0) {   /* up */
            const int row = y0 - istep;
            for (int col = smin_col; col <= smax_col; col++) {
                int i = row * width + col;
                if (!IN225(i) || !(sel[i] && (convex[i] || first))) continue;
                MARK(i - width);
                int j = i - width;
                if (convex[j]) {
                    if (col <= x0 && col > min_col) MARK(j - 1);
                    if (col >= x0 && col < max_col) MARK(j + 1);
                }
            }
        }
        if (y0 + istep < max_row) {   /* down */
            const int row = y0 + istep;
            for (int col = smin_col; col <= smax_col; col++) {
                int i = row * width + col;
                if (!IN225(i) || !(sel[i] && (convex[i] || first))) continue;
                MARK(i + width);
                int j = i + width;
                if (convex[j]) {
                    if (col <= x0 && col > min_col) MARK(j - 1);
                    if (col >= x0 && col < max_col) MARK(j + 1);
                }
            }
        }
        if (istep < x0) {   /* left */
            const int col = x0 - istep;
            for (int row = smin_row; row <= smax_row; row++) {
                int i = row * width + col;
                if (!IN225(i) || !(sel[i] && (convex[i] || first))) continue;
                MARK(i - 1);
                int j = i - 1;
                if (convex[j]) {
                    if (row <= y0 && row > min_row) MARK(j - width);
                    if (row >= y0 && row < max_row) MARK(j + width);
                }
            }
        }
        if (x0 + istep < max_col) {   /* right */
            const int col = x0 + istep;
            for (int row = smin_row; row <= smax_row; row++) {
                int i = row * width + col;
                if (!IN225(i) || !(sel[i] && (convex[i] || first))) continue;
                MARK(i + 1);
                int j = i + 1;
                if (convex[j]) {
                    if (row <= y0 && row > min_row) MARK(j - width);
                    if (row >= y0 && row < max_row) MARK(j + width);
                }
            }
        }
        if (num_new == 0) break;
        memcpy(sel, sel_new, sizeof(sel));
    }
#undef MARK
#undef IN225
    for (int i = 0; i < 225; i++) weights[i] = sel[i] ? (convex[i] ? 1.0 : nc_cost) : 0.0;
    if (add3x3)
        for (int row = y0 - 1; row <= y0 + 1; row++)
            if (row >= 0 && row < width)
                for (int col = x0 - 1; col <= x0 + 1; col++)
                    if (col >= 0 && col < width && weights[row * width + col] == 0.0) weights[row * width + col] = nc_cost;
}

typedef struct {
    int ns;
    double x[LMA_MAX_SAMPLES], y[LMA_MAX_SAMPLES], v[LMA_MAX_SAMPLES], w[LMA_MAX_SAMPLES];
    int si[LMA_MAX_SAMPLES], gi[LMA_MAX_SAMPLES];
    int ngroups, gnum[LMA_MAX_GROUPS], gdiag[LMA_MAX_GROUPS];   /* numbering by first sample */
    int input_diag;
    double scales[3];
    double all_pars[LMA_MAX_PARS];
    int par_mask[LMA_MAX_PARS], npars;
    double vector[LMA_MAX_PARS];
    int nvec;
    double group_weights[LMA_MAX_GROUPS], weights[LMA_MAX_POINTS], values[LMA_MAX_POINTS], pure_weight;
    int have_last;
    double last_rms[2], initial_rms[2];
    double last_ymfx[LMA_MAX_POINTS], last_jt[LMA_MAX_PARS][LMA_MAX_POINTS];
} lma_t;

/* Correlations2dLMA:196-210 addSample */
static void lma_add_sample(lma_t* L, double x, double y, double v, double w, int si, int gi) {
    int n = L->ns++;
    L->x[n] = x; L->y[n] = y; L->v[n] = v; L->w[n] = w; L->si[n] = si; L->gi[n] = gi;
    if (L->gnum[gi] < 0) {
        L->gdiag[L->ngroups] = L->input_diag;
        L->gnum[gi] = L->ngroups++;
    }
}

/* Correlations2dLMA:245-278 setWeightsValues */
static void lma_set_weights_values(lma_t* L, double half_width, double cost_wm, double cost_wxy) {
    const int np = L->ns;
    for (int g = 0; g < L->ngroups; g++) L->group_weights[g] = 0;
    L->weights[np] = cost_wm;
    L->weights[np + 1] = cost_wxy;
    L->values[np] = half_width;
    L->values[np + 1] = 0.0;
    double sw = 0;
    for (int i = 0; i < np; i++) {
        L->weights[i] = L->w[i];
        L->values[i] = L->v[i];
        L->group_weights[L->gnum[L->gi[i]]] += L->w[i];
        if (isnan(L->values[i]) || isnan(L->weights[i])) {
            L->weights[i] = 0.0;
            L->values[i] = 0.0;
        }
        sw += L->weights[i];
    }
    L->pure_weight = sw;
    sw += L->weights[np] + L->weights[np + 1];
    if (sw != 0.0) {
        sw = 1.0 / sw;
        for (int i = 0; i < np + 2; i++) L->weights[i] *= sw;
    }
    if (L->pure_weight > 0.0)
        for (int g = 0; g < L->ngroups; g++) L->group_weights[g] /= L->pure_weight;
    L->pure_weight *= sw;
}

/* Correlations2dLMA:212-243 initVector (+ toVector, 280-286) */
static void lma_init_vector(lma_t* L, const clt_corr_lma_params* lp, double x0) {
    L->npars = LMA_AG + L->ngroups;
    L->all_pars[LMA_X0] = x0;
    L->all_pars[LMA_WM] = lp->lma_half_width;
    L->all_pars[LMA_WYD] = 0.0;
    L->all_pars[LMA_WXY] = 0.0;
    L->par_mask[LMA_X0] = 1;
    L->par_mask[LMA_WM] = lp->lma_adjust_wm != 0;
    L->par_mask[LMA_WYD] = lp->lma_adjust_wy != 0;
    L->par_mask[LMA_WXY] = lp->lma_adjust_wxy != 0;
    for (int g = 0; g < L->ngroups; g++) L->par_mask[LMA_AG + g] = lp->lma_adjust_ag != 0;
    lma_set_weights_values(L, lp->lma_half_width, lp->lma_cost_wy, lp->lma_cost_wxy);
    int imx = 0;
    for (int i = 1; i < L->ns; i++)
        if (L->values[i] > L->values[imx]) imx = i;
    for (int g = 0; g < L->ngroups; g++) L->all_pars[LMA_AG + g] = L->values[imx];
    L->nvec = 0;
    for (int i = 0; i < L->npars; i++)
        if (L->par_mask[i]) L->vector[L->nvec++] = L->all_pars[i];
}

/* Correlations2dLMA:368-416 getFxJt: model Ag * (1 - (dx/Wx)^2 - (y/Wy)^2) (x sqrt2 on diagonals),
 * Wy = Wm * scale + Wyd, Wx = Wy + Wxy; two regularisation points Wm -> half_width, Wxy -> 0. */
static void lma_fx_jt(const lma_t* L, const double* vec, double* fx, double (*jt)[LMA_MAX_POINTS]) {
    double av[LMA_MAX_PARS];
    for (int i = 0, k = 0; i < L->npars; i++) av[i] = L->par_mask[i] ? vec[k++] : L->all_pars[i];
    const double sqrt2 = sqrt(2.0);
    const int ns = L->ns;
    for (int n = 0; n < ns; n++) {
        const int grp = L->gnum[L->gi[n]];
        const double wscale = L->gdiag[grp] ? sqrt2 : 1.0;
        const double Wy = av[LMA_WM] * L->scales[L->si[n]] + av[LMA_WYD];
        const double Wx = Wy + av[LMA_WXY];
        const double dx = L->x[n] - av[LMA_X0];
        const double Ag = av[LMA_AG + grp];
        const double dxw = wscale * dx / Wx;
        const double dyw = wscale * L->y[n] / Wy;
        const double d = (1.0 - dxw * dxw - dyw * dyw);
        fx[n] = d * Ag;
        if (jt) {
            int np = 0;
            if (L->par_mask[LMA_X0]) jt[np++][n] = wscale * Ag * 2 * dxw / Wx;
            const double dfdWx = Ag * 2 * dxw * dxw / Wx;
            const double dfdWy = Ag * 2 * dyw * dyw / Wy;
            if (L->par_mask[LMA_WM]) jt[np++][n] = L->scales[L->si[n]] * (dfdWx + dfdWy);
            if (L->par_mask[LMA_WYD]) jt[np++][n] = (dfdWx + dfdWy);
            if (L->par_mask[LMA_WXY]) jt[np++][n] = dfdWx;
            for (int g = 0; g < L->ngroups; g++)
                if (L->par_mask[LMA_AG + g]) jt[np++][n] = (g == grp) ? d : 0.0;
        }
    }
    fx[ns] = av[LMA_WM];
    fx[ns + 1] = av[LMA_WXY];
    if (jt)
        for (int i = 0, np = 0; i < L->npars; i++)
            if (L->par_mask[i]) {
                jt[np][ns] = (i == LMA_WM) ? 1.0 : 0.0;
                jt[np++][ns + 1] = (i == LMA_WXY) ? 1.0 : 0.0;
            }
}

/* Correlations2dLMA:442-461 getWYmFxRms: fx <- w * (y - fx), returns {rms, rms_pure} */
static void lma_ymfx_rms(const lma_t* L, double* fx, double rms_out[2]) {
    double rms = 0;
    for (int i = 0; i < L->ns; i++) {
        double d = L->values[i] - fx[i];
        fx[i] = L->weights[i] * d;
        rms += fx[i] * d;
    }
    const double rms_pure = sqrt(rms) / L->pure_weight;
    for (int i = L->ns; i < L->ns + 2; i++) {
        double d = L->values[i] - fx[i];
        fx[i] = L->weights[i] * d;
        rms += fx[i] * d;
    }
    rms_out[0] = sqrt(rms);
    rms_out[1] = rms_pure;
}

/* Inverse of a small dense matrix by LU decomposition with partial pivoting (the Jama
 * Matrix.inverse() of Correlations2dLMA:556); returns 0 when a pivot is exactly zero. */
static int small_inverse(int n, double a[LMA_MAX_PARS][LMA_MAX_PARS], double inv[LMA_MAX_PARS][LMA_MAX_PARS]) {
    int piv[LMA_MAX_PARS];
    for (int i = 0; i < n; i++) piv[i] = i;
    for (int j = 0; j < n; j++) {
        double col[LMA_MAX_PARS];
        for (int i = 0; i < n; i++) col[i] = a[i][j];
        for (int i = 0; i < n; i++) {
            int kmax = (i < j) ? i : j;
            double s = 0.0;
            for (int k = 0; k < kmax; k++) s += a[i][k] * col[k];
            a[i][j] = col[i] -= s;
        }
        int p = j;
        for (int i = j + 1; i < n; i++)
            if (fabs(col[i]) > fabs(col[p])) p = i;
        if (p != j) {
            for (int k = 0; k < n; k++) {
                double t = a[p][k];
                a[p][k] = a[j][k];
                a[j][k] = t;
            }
            int t = piv[p];
            piv[p] = piv[j];
            piv[j] = t;
        }
        if (a[j][j] != 0.0)
            for (int i = j + 1; i < n; i++) a[i][j] /= a[j][j];
    }
    for (int j = 0; j < n; j++)
        if (a[j][j] == 0.0) return 0;
    for (int c = 0; c < n; c++) {   /* solve for column c of the identity */
        double x[LMA_MAX_PARS];
        for (int i = 0; i < n; i++) x[i] = (piv[i] == c) ? 1.0 : 0.0;
        for (int k = 0; k < n; k++)
            for (int i = k + 1; i < n; i++) x[i] -= x[k] * a[i][k];
        for (int k = n - 1; k >= 0; k--) {
            x[k] /= a[k][k];
            for (int i = 0; i < k; i++) x[i] -= x[k] * a[i][k];
        }
        for (int i = 0; i < n; i++) inv[i][c] = x[i];
    }
    return 1;
}

/* Correlations2dLMA:533-640 lmaStep; returns success in ok[0], done in ok[1] */
static void lma_step(lma_t* L, double lambda, double rms_diff, int ok[2]) {
    const int npts = L->ns + 2, nv = L->nvec;
    ok[0] = ok[1] = 0;
    if (!L->have_last) {
        lma_fx_jt(L, L->vector, L->last_ymfx, L->last_jt);
        lma_ymfx_rms(L, L->last_ymfx, L->last_rms);
        L->initial_rms[0] = L->last_rms[0];
        L->initial_rms[1] = L->last_rms[1];
        L->have_last = 1;
    }
    double m[LMA_MAX_PARS][LMA_MAX_PARS], inv[LMA_MAX_PARS][LMA_MAX_PARS];
    for (int i = 0; i < nv; i++)   /* Correlations2dLMA:418-440 getWJtJlambda */
        for (int j = i; j < nv; j++) {
            double d = 0.0;
            for (int k = 0; k < npts; k++) d += L->weights[k] * L->last_jt[i][k] * L->last_jt[j][k];
            m[i][j] = d;
            if (i == j) m[i][j] += d * lambda;
            else m[j][i] = d;
        }
    if (!small_inverse(nv, m, inv)) {
        ok[1] = 1;
        return;
    }
    double jty[LMA_MAX_PARS], nvec[LMA_MAX_PARS];
    for (int i = 0; i < nv; i++) {
        double s = 0;
        for (int k = 0; k < npts; k++) s += L->last_jt[i][k] * L->last_ymfx[k];
        jty[i] = s;
    }
    for (int i = 0; i < nv; i++) {
        double s = 0;
        for (int k = 0; k < nv; k++) s += inv[i][k] * jty[k];
        nvec[i] = L->vector[i] + s;
    }
    lma_fx_jt(L, nvec, L->last_ymfx, L->last_jt);
    double rms[2];
    lma_ymfx_rms(L, L->last_ymfx, rms);
    if (rms[0] < L->last_rms[0]) {
        ok[0] = 1;
        ok[1] = rms[0] >= (L->last_rms[0] * (1.0 - rms_diff));
        L->last_rms[0] = rms[0];
        L->last_rms[1] = rms[1];
        for (int i = 0; i < nv; i++) L->vector[i] = nvec[i];
    } else {
        lma_fx_jt(L, L->vector, L->last_ymfx, L->last_jt);
        lma_ymfx_rms(L, L->last_ymfx, L->last_rms);
    }
}

/* Correlations2dLMA:477-530 runLma */
static int lma_run(lma_t* L, const clt_corr_lma_params* lp) {
    double lambda = lp->lma_lambda_initial;
    int ok[2] = {0, 0};
    for (int iter = 0; iter < lp->lma_num_iter; iter++) {
        lma_step(L, lambda, lp->lma_rms_diff, ok);
        if (ok[1]) break;
        if (ok[0]) lambda *= lp->lma_lambda_scale_good;
        else {
            lambda *= lp->lma_lambda_scale_bad;
            if (lambda > lp->lma_lambda_max) break;
        }
    }
    if (ok[0]) return 1;
    return L->last_rms[0] < L->initial_rms[0];
}

/* C2:1754-1820 / C2:1903-1935 the samples of corrLMA: the compatible pair groups combined
 * (C2:462-488), filterConvex around the start point and addSamples (C2:2230-2338). */
static void lma_samples(lma_t* L, const double (*corrs)[225], const clt_pair* pairs, int npairs, const char* sel,
                        double xcenter, const clt_corr_lma_params* lp) {
    L->ns = L->ngroups = L->have_last = L->input_diag = 0;
    for (int g = 0; g < LMA_MAX_GROUPS; g++) L->gnum[g] = -1;
    int sl[3], nsl = 0, g_num[6], g_scale_ind[6];
    for (int i = 0; i < 6; i++) {   /* getNumberBaseOfCompatiblePairs (C2:502-519) */
        g_num[i] = 0;
        for (int p = 0; p < npairs; p++)
            if (sel[p] && pair_compatible(&pairs[p], LMA_GROUPS[i][0], LMA_GROUPS[i][1])) g_num[i]++;
        if (g_num[i] > 0) {
            int k = 0;
            while (k < nsl && sl[k] != LMA_GROUPS[i][1]) k++;
            if (k == nsl) sl[nsl++] = LMA_GROUPS[i][1];
            g_scale_ind[i] = k;
        }
    }
    for (int k = 0; k < nsl; k++) L->scales[k] = sl[k];
    const int hwin = lp->cnvx_hwnd_size, hw2 = 2 * hwin + 1, center = 7, width = 15;
    const int center_index = (width + 1) * center;
    static const int quad_signs[4][2] = {{-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
    int ig = 0;
    for (int i = 0; i < 6; i++) {
        if (g_num[i] == 0) continue;
        const int diag = LMA_GROUPS[i][0];
        const double scale = sl[g_scale_ind[i]], scale05 = scale / 2.0;
        double combo[225], filt[225];
        combine_compatible(corrs, pairs, npairs, sel, diag, LMA_GROUPS[i][1], combo);
        const int ixc = (int)jround(xcenter / scale);
        const double xcs = ixc * scale;
        filter_convex(combo, hwin, ixc, diag ? ixc : 0, lp->cnvx_add3x3, lp->cnvx_weight, filt);
        L->input_diag = diag;
        for (int arow = 0; arow < hw2; arow += diag ? 1 : 2) {
            for (int acol = diag ? (arow & 1) : 0; acol < hw2; acol += 2) {
                for (int quad = 0; quad < 4; quad++) {
                    if (!((arow > 0) || (quad & 2)) || !((acol > 0) || (quad & 1))) continue;
                    const int sx = quad_signs[quad][0], sy = quad_signs[quad][1];
                    const int cx = diag ? (sx * acol - sy * arow) / 2 + ixc : (sx * acol) / 2 + ixc;
                    const int cy = diag ? (sx * acol + sy * arow) / 2 + ixc : (sy * arow) / 2;
                    if (cx < -center || cx > center || cy < -center || cy > center) continue;
                    double w = filt[center_index + width * cy + cx] * g_num[i];
                    if (!(w > 0.0)) continue;
                    const double v = combo[center_index + width * cy + cx];
                    if (lp->ortho_vasw_pwr != 0) w *= pow(fabs(v), lp->ortho_vasw_pwr);
                    lma_add_sample(L, sx * acol * scale05 + xcs, sy * arow * scale05, v, w, g_scale_ind[i], ig);
                }
            }
        }
        ig++;
    }
}

/* C2:1754-1900 corrLMA (run_poly_instead = false); on success out = {disparity, strength, Wm, Wxy}
 * (getDisparityStrengthWidth, Correlations2dLMA:165-185). */
static int corr_lma(const double (*corrs)[225], const clt_pair* pairs, int npairs, const char* sel, double xcenter,
                    const clt_corr_lma_params* lp, double out[4]) {
    lma_t* L = malloc(sizeof(lma_t));
    lma_samples(L, corrs, pairs, npairs, sel, xcenter, lp);
    lma_init_vector(L, lp, xcenter);
    const int ok = lma_run(L, lp);
    for (int i = 0, k = 0; i < L->npars; i++)   /* updateFromVector */
        if (L->par_mask[i]) L->all_pars[i] = L->vector[k++];
    if (ok) {
        double sum_amp = 0.0, max_amp = 0.0;
        for (int g = 0; g < L->ngroups; g++) sum_amp += L->group_weights[g] * L->all_pars[LMA_AG + g];
        for (int n = 0; n < L->ns; n++)
            if (L->v[n] > max_amp) max_amp = L->v[n];
        if (sum_amp > 1.25 * max_amp) sum_amp = max_amp;
        out[0] = -L->all_pars[LMA_X0];
        out[1] = sum_amp;
        out[2] = L->all_pars[LMA_WM];
        out[3] = L->all_pars[LMA_WXY];
    }
    free(L);
    return ok;
}

/* C2:1663-1750 foregroundCorrect: the disparity to use (full_disp unless corrected) */
static double foreground_correct(const clt_corr_lma_params* lp, const int have[4], double dds[4][4], double full_disp) {
    const int bg = lp->fo_far != 0, ortho = lp->fo_ortho != 0;
    double mn = NAN, mx = NAN;
    int imn = -1, imx = -1;
    double eff[4] = {0, 0, 0, 0}, width_d[4] = {0, 0, 0, 0};
    int strong[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        if (have[i]) {
            if (isnan(mn) || dds[i][0] < mn) { mn = dds[i][0]; imn = i; }
            if (isnan(mx) || dds[i][0] > mx) { mx = dds[i][0]; imx = i; }
            eff[i] = dds[i][1] * (dds[i][2] - dds[i][3]) / (dds[i][2] + dds[i][3]);
            width_d[i] = dds[i][2] + dds[i][3];
            strong[i] = (dds[i][1] >= lp->fo_min_strength) && (eff[i] >= lp->fo_min_eff);
        } else {
            eff[i] = NAN;
        }
    }
    const int are_ortho = (imn ^ imx) == 1;
    double corr = lp->fo_overcorrection * (mx - mn);
    const int isel = bg ? imn : imx, iother = bg ? imx : imn;
    if (isel < 0 || !strong[isel]) corr = NAN;
    else if (width_d[isel] > lp->fo_max_hwidth) corr = NAN;
    else if (eff[isel] / eff[isel ^ 1] < lp->fo_min_eff_ratio) corr = NAN;
    else {
        if (ortho && !are_ortho) corr = 0.0;
        else if (!strong[iother]) corr = 0.0;
        if (!(mx >= (mn + lp->fo_min_diff))) corr = NAN;
    }
    double disp = full_disp;
    if (!isnan(corr)) {
        if (bg) {
            double lim = full_disp - (full_disp - mn) * lp->fo_lim_overcorr;
            disp = fmax(mn - corr, lim);
        } else {
            double lim = full_disp + (mx - full_disp) * lp->fo_lim_overcorr;
            disp = fmin(mx + corr, lim);
        }
    }
    return disp;
}

/* ID:2254-2336 for one tile: the LMA over all oriented pairs (C2:1754), mixing into CM/strength
 * (ID:2273-2281), the four one-direction fits and the foreground correction (ID:2288-2336).
 * disp: the tile's record; reads CLT_DISP_CM_RAW / CLT_DISP_STR_RAW, writes CM, STRENGTH, POLY. */
static void lma_chain(const double (*corrs)[225], const clt_pair* pairs, int npairs, const clt_corr_lma_params* lp,
                      double* disp) {
    disp[CLT_DISP_POLY] = disp[CLT_DISP_POLY_STR] = NAN;
    const double cm = disp[CLT_DISP_CM_RAW];
    double strength = disp[CLT_DISP_STR_RAW];
    disp[CLT_DISP_CM] = cm;
    disp[CLT_DISP_STRENGTH] = strength;
    if (!lp->enable || isnan(cm) || !(strength > lp->min_poly_strength)) return;
    char sel[CLT_MAX_PAIRS];
    for (int p = 0; p < npairs; p++) sel[p] = pairs[p].dir >= 0;
    double r[4];
    if (!corr_lma(corrs, pairs, npairs, sel, -cm, lp, r)) return;
    double disparity = cm;
    disp[CLT_DISP_POLY] = r[0];
    disp[CLT_DISP_POLY_STR] = r[1];
    if (lp->mix_corr_poly) {
        disparity = r[0];
        strength = r[1];
        disp[CLT_DISP_CM] = disparity;
        disp[CLT_DISP_STRENGTH] = strength;
    }
    if (!(lp->fo_correct && strength > 0)) return;
    double dds[4][4];
    int have[4];
    for (int dir = 0; dir < 4; dir++) {   /* C2:1591-1630 corr4dirsLMA */
        int any = 0;
        for (int p = 0; p < npairs; p++) any |= (sel[p] = (pairs[p].dir == dir));
        have[dir] = any && corr_lma(corrs, pairs, npairs, sel, -disparity, lp, dds[dir]);
    }
    const double mod = foreground_correct(lp, have, dds, disparity);
    if (mod != disparity && lp->fo_correct && strength > lp->fo_min_strength) disp[CLT_DISP_CM] = mod;
}

typedef struct {
    frame_t f;
    const double* clt;
    const int32_t* tasks;
    const clt_pair* pairs;
    int npairs;
    clt_corr_params prm;
    clt_corr_lma_params lma;
    real lpf[64];
    real col_weights[3];
    double wndy[16], wndx[16], notch[15];
    int wndy_len, wndx_len;
    double* corr_out;   /* [task][pair][225] */
    double* combo_out;  /* [task][strip_height*15] or NULL */
    double* disp_out;   /* [task][CLT_DISP_LAYERS] or NULL */
} corr_ctx;

/* ID:2101-2238 for one task tile */
static void corr_task(void* vctx, int t) {
    corr_ctx* c = (corr_ctx*)vctx;
    const clt_geom* g = c->f.g;
    const int K = g->num_colors, hw = c->prm.strip_height, slen = hw * 15;
    double* corr_o = c->corr_out + (size_t)t * c->npairs * 225;
    double* combo_o = c->combo_out ? c->combo_out + (size_t)t * slen : NULL;
    double* disp_o = c->disp_out ? c->disp_out + (size_t)t * CLT_DISP_LAYERS : NULL;
    if (disp_o) {   /* ID:1713-1719 init: NaN, strengths 0 */
        for (int i = 0; i < CLT_DISP_LAYERS; i++) disp_o[i] = NAN;
        disp_o[CLT_DISP_STRENGTH] = disp_o[CLT_DISP_HOR_STR] = disp_o[CLT_DISP_VERT_STR] = 0;
        disp_o[CLT_DISP_CM_WEIGHT] = disp_o[CLT_DISP_STR_RAW] = 0;
    }
    if (combo_o)
        for (int i = 0; i < slen; i++) combo_o[i] = NAN;
    if (task_flags(g, c->tasks, t) == 0) {
        for (int i = 0; i < c->npairs * 225; i++) corr_o[i] = NAN;
        return;
    }
    int tx, ty;
    task_txy(g, c->tasks, t, &tx, &ty);
    size_t tile = (size_t)ty * c->f.tiles_x + tx;
    real corrs[CLT_MAX_PAIRS][225];
    real (*clt1)[4][64] = malloc(sizeof(real) * 3 * 256);
    real (*clt2)[4][64] = malloc(sizeof(real) * 3 * 256);
    for (int p = 0; p < c->npairs; p++) {
        for (int col = 0; col < K; col++) {
            load_clt(&c->f, c->clt, c->pairs[p].cam1, tile, col, clt1[col]);
            load_clt(&c->f, c->clt, c->pairs[p].cam2, tile, col, clt2[col]);
        }
        correlate_composite_fd(K, clt1, clt2, c->lpf, c->col_weights, (real)c->prm.fat_zero, corrs[p]);
        for (int i = 0; i < 225; i++) corr_o[p * 225 + i] = corrs[p][i];
    }
    free(clt1);
    free(clt2);
    /* strips (C2:556-578): pairs with a known orientation; ss relative to the smallest scale */
    int ss_min = 1 << 30;
    for (int p = 0; p < c->npairs; p++)
        if (c->pairs[p].dir >= 0 && c->pairs[p].scale < ss_min) ss_min = c->pairs[p].scale;
    real* strips[CLT_MAX_PAIRS];
    real strip_buf[CLT_MAX_PAIRS][16 * 15];
    int nstrip = 0;
    for (int p = 0; p < c->npairs; p++) {
        strips[p] = NULL;
        if (c->pairs[p].dir < 0) continue;
        strip_single(corrs[p], hw, c->pairs[p].dir, c->pairs[p].scale / ss_min, strip_buf[p]);
        strips[p] = strip_buf[p];
        nstrip++;
    }
    if (nstrip == 0) return;
    real combo[16 * 15];
    combine_strips(strips, c->npairs, slen, (real)c->prm.corr_offset, c->prm.twice_diagonal, combo);
    if (combo_o)
        for (int i = 0; i < slen; i++) combo_o[i] = combo[i];
    if (!disp_o) return;
    int ix;
    if (max_xy_int_axis(combo, (real)c->prm.min_corr, &ix)) {   /* ID:2186-2208 */
        disp_o[CLT_DISP_INT] = -ix;
        disp_o[CLT_DISP_STRENGTH] = combo[ix + 7];
        real cm[3];
        if (max_x_cm(combo, hw, ix, c->wndy, c->wndy_len, c->wndx, c->wndx_len, cm)) {
            disp_o[CLT_DISP_CM] = -cm[0];
            disp_o[CLT_DISP_CM_WEIGHT] = cm[1];
        }
    }
    for (int vert = 0; vert < 2; vert++) {   /* ID:2211-2238, getMaskHorizontal(1)/getMaskVertical(1) */
        real* sel[CLT_MAX_PAIRS];
        int n = 0;
        for (int p = 0; p < c->npairs; p++)
            if (c->pairs[p].dir == vert && c->pairs[p].scale == 1) sel[n++] = corrs[p];
        real r[2];
        if (max_xs_ortho(sel, n, (real)c->prm.corr_offset, vert, c->notch, r)) {
            disp_o[vert ? CLT_DISP_VERT : CLT_DISP_HOR] = -r[0];
            disp_o[vert ? CLT_DISP_VERT_STR : CLT_DISP_HOR_STR] = r[1];
        }
    }
    disp_o[CLT_DISP_CM_RAW] = disp_o[CLT_DISP_CM];
    disp_o[CLT_DISP_STR_RAW] = disp_o[CLT_DISP_STRENGTH];
#ifdef ORC_FP32
    static __thread double dcorrs[CLT_MAX_PAIRS][225];
    for (int p = 0; p < c->npairs; p++)
        for (int i = 0; i < 225; i++) dcorrs[p][i] = corrs[p][i];
    lma_chain((const double (*)[225])dcorrs, c->pairs, c->npairs, &c->lma, disp_o);
#else
    lma_chain((const double (*)[225])corrs, c->pairs, c->npairs, &c->lma, disp_o);
#endif
}

static void color_weights(const clt_geom* g, double red, double blue, real* w) {
    if (g->num_colors == 1) {
        w[0] = 1;
        return;
    }
    double gw = 1.0 / (1.0 + red + blue);   /* ID:1571-1581 */
    w[2] = (real)gw;
    w[0] = (real)(red * gw);
    w[1] = (real)(blue * gw);
}

/* ================================ IMCLT (ID:3559-3666) ================================ */
/* 16x16 inverse of one camera/colour tile, already LPF-ed/scaled: sum over modes */
static void iclt_tile(real clt[4][64], real scale, const real* lpf, real* out256) {
    real tmp[64], idct[64], unf[256];
    for (int i = 0; i < 256; i++) out256[i] = 0;
    for (int m = 0; m < 4; m++) {
        for (int i = 0; i < 64; i++) tmp[i] = clt[m][i] * (lpf ? scale * lpf[i] : scale);
        int idct_mode = ((m << 1) & 2) | ((m >> 1) & 1);
        dttt_iv(tmp, idct_mode, idct);
        unfold_tile(idct, m, unf);
        for (int i = 0; i < 256; i++) out256[i] += unf[i];
    }
}

typedef struct {
    frame_t f;
    const double* clt;
    const uint8_t* valid;   /* [tile] */
    const real* lpf;        /* NULL: none */
    int cam, color, pass;
    real* plane;            /* (H+8) x (W+8) */
} imclt_ctx;

static void imclt_tile_fn(void* vctx, int n) {
    imclt_ctx* c = (imclt_ctx*)vctx;
    int nx = (c->f.tiles_x + 1 - (c->pass & 1)) / 2;
    int tx = 2 * (n % nx) + (c->pass & 1), ty = 2 * (n / nx) + ((c->pass >> 1) & 1);
    size_t tile = (size_t)ty * c->f.tiles_x + tx;
    if (!c->valid[tile]) return;
    real clt[4][64], out[256];
    load_clt(&c->f, c->clt, c->cam, tile, c->color, clt);
    iclt_tile(clt, (real)0.25, c->lpf, out);   /* debug_scale = 1/4 for mask 15 (ID:3574) */
    const int ow = c->f.g->width + 8;
    for (int i = 0; i < 16; i++)
        for (int j = 0; j < 16; j++) c->plane[(size_t)(ty * 8 + i) * ow + tx * 8 + j] += out[16 * i + j];
}

/* ================================ textures (ID:2662-2812, 2830-3147) ================================ */
static const double KERN_G[9] = {0.0, 0.125, 0.0, 0.125, 0.5, 0.125, 0.0, 0.125, 0.0};          /* ID:43-46 */
static const double KERN_RB[9] = {0.0625, 0.125, 0.0625, 0.125, 0.25, 0.125, 0.0625, 0.125, 0.0625}; /* ID:47-50 */
static const int CORN_SIDE[9][9] = {                                                                /* ID:52-61 */
    {4, 5, 7, 8, -1}, {3, 4, 5, 6, 7, 8, -1}, {3, 4, 6, 7, -1}, {1, 2, 4, 5, 7, 8, -1},
    {0, 1, 2, 3, 4, 5, 6, 7, 8}, {0, 1, 3, 4, 6, 7, -1}, {1, 2, 4, 5, -1}, {0, 1, 2, 3, 4, 5, -1},
    {0, 1, 3, 4, -1}};

/* ID:3048-3147 tile_debayer (16x16) */
static void tile_debayer(int rb, const real* tile, real* rslt) {
    const int ts = 16, tsm1 = 15;
    const int neib[9] = {-ts - 1, -ts, -ts + 1, -1, 0, 1, ts - 1, ts, ts + 1};
    const double* kern = rb ? KERN_RB : KERN_G;
    real k_corn = rb ? (real)(16.0 / 9.0) : (real)(4.0 / 3.0);
    real k_side = rb ? (real)(4.0 / 3.0) : (real)(8.0 / 7.0);
    for (int i = 0; i < 256; i++) rslt[i] = 0;
    for (int i = 0; i < ts; i++) {
        for (int j = 0; j < ts; j++) {
            int side = ((i == 0) ? 0 : (i == tsm1) ? 6 : 3) + ((j == 0) ? 0 : (j == tsm1) ? 2 : 1);
            int indx = i * ts + j;
            real s = 0;
            for (int k = 0; k < 9 && CORN_SIDE[side][k] >= 0; k++) {
                int dr = CORN_SIDE[side][k];
                s += tile[indx + neib[dr]] * (real)kern[dr];
            }
            if (side == 4) rslt[indx] = s;
            else if (side == 0 || side == 2 || side == 6 || side == 8) rslt[indx] = s * k_corn;
            else rslt[indx] = s * k_side;
        }
    }
}

/* ID:3022-3045 tile_debayer_shot_corr */
static void tile_debayer_shot_corr(int rb, const real* tile, const real* window2, real min_shot,
                                   real scale_shot, real* out) {
    real nw[256];
    for (int i = 0; i < 256; i++) nw[i] = tile[i] / window2[i];
    tile_debayer(rb, nw, out);
    if (scale_shot > 0) {
        real k = (real)(1.0 / sqrt((double)min_shot));
        for (int i = 0; i < 256; i++)
            out[i] = scale_shot * ((out[i] > min_shot) ? (real)sqrt((double)out[i]) : (k * out[i]));
    }
    for (int i = 0; i < 256; i++) out[i] = out[i] * window2[i];
}

/* ID:3002-3019 dust_remove */
static void dust_remove(wreal pw[][256], int np) {
    for (int i = 0; i < 256; i++) {
        int wi = 0;
        for (int ip = 1; ip < np; ip++)
            if (pw[ip][i] < pw[wi][i]) wi = ip;
        wreal avg = 0;
        for (int ip = 1; ip < np; ip++)
            if (ip != wi) avg += pw[ip][i];
        avg /= (np - 1);
        wreal scale = 1 + (avg - pw[wi][i]) / (avg * (np - 1));
        for (int ip = 1; ip < np; ip++)
            if (ip != wi) pw[ip][i] *= scale;
        pw[wi][i] *= pw[wi][i] / avg;
    }
}

typedef struct {
    frame_t f;
    const double* clt;
    const int32_t* tasks;
    const double* extra_disparity;   /* [task] or NULL */
    clt_tex_params prm;
    real lpf[64];
    real col_weights[3];
    real window2[256];
    int planes;
    double* tiles_out;   /* [task][planes][256] */
    double* stats_out;   /* [task][4*C] or NULL */
} tex_ctx;

/* tile_combine_rgba (ID:2830-3000) with keep_weights, on the debayered tiles deb[port][colour][256],
 * then the final mix of the tiles iclt[port][colour][256] with the port weights (ID:2789-2801):
 * out [planes][256] (RGBA, then with planes > 4 one weight per port and the rms of r, b, g and the
 * weighted rms), stats [4 * ports] (max_diff, then the mean tones r, b, g of every port) or NULL. */
static void combine_rgba(int ports, int port_mask, real (*deb)[3][256], real (*iclt)[3][256], const real* lt_window,
                         const real* chw, const clt_tex_params* prm, int planes, double* out, double* stats) {
    const int numcol = 3;
    int used = 0;
    for (int ip = 0; ip < ports; ip++) used += (port_mask >> ip) & 1;
    wreal (*pw)[256] = calloc((size_t)ports, sizeof(wreal) * 256);
    real avg[3][256], alpha[256], rms[4][256];
    for (int i = 0; i < 256; i++) {   /* rms for keep_weights (ID:2858-2874) */
        real sw = 0;
        for (int chn = 0; chn < numcol; chn++) {
            real s0 = 0, s1 = 0, s2 = 0;
            for (int ip = 0; ip < ports; ip++)
                if ((port_mask >> ip) & 1) {
                    s0 += 1;
                    s1 += deb[ip][chn][i];
                    s2 += deb[ip][chn][i] * deb[ip][chn][i];
                }
#ifdef ORC_FP32
            /* the same standard deviation in two passes: s0 s2 - s1^2 cancels in float */
            real m = s1 / s0, v = 0;
            for (int ip = 0; ip < ports; ip++)
                if ((port_mask >> ip) & 1) v += (deb[ip][chn][i] - m) * (deb[ip][chn][i] - m);
            rms[chn][i] = (real)sqrt((double)(v / s0));
            (void)s2;
#else
            rms[chn][i] = (real)sqrt((double)(s0 * s2 - s1 * s1)) / s0;
#endif
            sw += chw[chn] * rms[chn][i] * rms[chn][i];
        }
        rms[3][i] = (real)sqrt((double)sw);
    }
    if (used > 1) {
        real ksigma = (real)(1.0 / (2.0 * prm->diff_sigma * prm->diff_sigma));
        int npairs = 0, pp[CLT_MAX_CAMS * CLT_MAX_CAMS / 2][2];
        real pd2r[CLT_MAX_CAMS * CLT_MAX_CAMS / 2];
        for (int i = 0; i < ports; i++)
            if ((port_mask >> i) & 1)
                for (int j = i + 1; j < ports; j++)
                    if ((port_mask >> j) & 1) {
                        real dx = prm->port_offsets[j][0] - prm->port_offsets[i][0];
                        real dy = prm->port_offsets[j][1] - prm->port_offsets[i][1];
                        pp[npairs][0] = i;
                        pp[npairs][1] = j;
                        pd2r[npairs++] = ksigma / (dx * dx + dy * dy);
                    }
        for (int i = 0; i < 256; i++) {
            for (int ip = 0; ip < ports; ip++) pw[ip][i] = 0;
            for (int ip = 0; ip < npairs; ip++) {
                real d = 0;
                for (int chn = 0; chn < numcol; chn++) {
                    real dc = deb[pp[ip][0]][chn][i] - deb[pp[ip][1]][chn][i];
                    dc /= lt_window[i];
                    d += chw[chn] * dc * dc;
                }
                const wreal e = exp((double)(-pd2r[ip] * d));
                pw[pp[ip][0]][i] += e;
                pw[pp[ip][1]][i] += e;
            }
            int b1 = 0;
            for (int ip = b1 + 1; ip < ports; ip++)
                if (pw[ip][i] > pw[b1][i]) b1 = ip;
            int b2 = (b1 == 0) ? 1 : 0;
            for (int ip = b2 + 1; ip < ports; ip++)
                if ((ip != b1) && (pw[ip][i] > pw[b2][i])) b2 = ip;
            real w1 = (real)(pw[b1][i] / (pw[b1][i] + pw[b2][i]));
            real w2 = 1 - w1;
            for (int chn = 0; chn < numcol; chn++) avg[chn][i] = w1 * deb[b1][chn][i] + w2 * deb[b2][chn][i];
            for (int ip = 0; ip < ports; ip++)
                if ((port_mask >> ip) & 1) {
                    real d2 = 0;
                    for (int chn = 0; chn < numcol; chn++) {
                        real dc = deb[ip][chn][i] - avg[chn][i];
                        dc /= lt_window[i];
                        d2 += chw[chn] * dc * dc;
                    }
                    pw[ip][i] = exp((double)(-ksigma * d2));
                }
            wreal k = 0;
            for (int ip = 0; ip < ports; ip++) k += pw[ip][i];
            k = 1 / k;
            for (int chn = 0; chn < numcol; chn++) {
                avg[chn][i] = 0;
                for (int ip = 0; ip < ports; ip++) avg[chn][i] += (real)(k * pw[ip][i]) * deb[ip][chn][i];
            }
        }
    } else if (used > 0) {
        for (int ip = 0; ip < ports; ip++)
            if ((port_mask >> ip) & 1)
                for (int i = 0; i < 256; i++) {
                    for (int chn = 0; chn < numcol; chn++) avg[chn][i] = deb[ip][chn][i];
                    pw[ip][i] = 1;
                }
    }
    if (prm->dust_remove && (used == 4)) dust_remove(pw, ports);
    for (int i = 0; i < 256; i++) {
        wreal a = 0;
        for (int ip = 0; ip < ports; ip++)
            if ((port_mask >> ip) & 1) a += pw[ip][i];
        alpha[i] = (real)(a * (lt_window[i] / used));   /* ID:2958: alpha *= lt_window / used */
    }
    if (stats) {   /* max_diff (ID:2965-2981), ports_rgb (ID:2982-2996) */
        for (int ip = 0; ip < ports; ip++) {
            real md = 0;
            if ((port_mask >> ip) & 1)
                for (int i = 0; i < 256; i++) {
                    real d2 = 0;
                    for (int chn = 0; chn < numcol; chn++) {
                        real dc = deb[ip][chn][i] - avg[chn][i];
                        d2 += dc * dc * chw[chn];
                    }
                    d2 *= lt_window[i];
                    if (d2 > md) md = d2;
                }
            stats[ip] = sqrt((double)md);
        }
        for (int ip = 0; ip < ports; ip++)
            for (int chn = 0; chn < numcol; chn++) {
                real s = 0;
                for (int i = 0; i < 256; i++) s += deb[ip][chn][i];
                stats[ports + chn * ports + ip] = s / 256;
            }
    }
    /* ID:2789-2801 final mix of the non-debayered tiles with the port weights */
    for (int i = 0; i < 256; i++) {
        wreal sw = 0;
        for (int ip = 0; ip < ports; ip++) sw += pw[ip][i];
        if (sw != 0) sw = 1 / sw;
        for (int chn = 0; chn < numcol; chn++) {
            real s = 0;
            for (int ip = 0; ip < ports; ip++) s += (real)(sw * pw[ip][i]) * iclt[ip][chn][i];
            out[chn * 256 + i] = s;
        }
        out[3 * 256 + i] = alpha[i];
        if (planes > 4) {
            for (int ip = 0; ip < ports; ip++) out[(4 + ip) * 256 + i] = pw[ip][i];
            for (int k = 0; k < 4; k++) out[(4 + ports + k) * 256 + i] = rms[k][i];
        }
    }
    free(pw);
}

/* ID:2659-2812 for one task tile, then tile_combine_rgba (ID:2830-3000) */
static void tex_task(void* vctx, int t) {
    tex_ctx* c = (tex_ctx*)vctx;
    const clt_geom* g = c->f.g;
    const int ports = g->num_cams, numcol = 3;
    double* out = c->tiles_out + (size_t)t * c->planes * 256;
    double* stats = c->stats_out ? c->stats_out + (size_t)t * 4 * ports : NULL;
    int32_t flags = task_flags(g, c->tasks, t);
    if (flags == 0) {
        for (int i = 0; i < c->planes * 256; i++) out[i] = NAN;
        if (stats)
            for (int i = 0; i < 4 * ports; i++) stats[i] = NAN;
        return;
    }
    int port_mask = flags & 0xff & ((1 << ports) - 1);
    int tx, ty;
    task_txy(g, c->tasks, t, &tx, &ty);
    size_t tile = (size_t)ty * c->f.tiles_x + tx;
    double extra = c->extra_disparity ? c->extra_disparity[t] : 0.0;
    if (isnan(extra)) extra = 0;
    int forced = (flags >> 24) & 1;   /* C2:97 getForcedDisparity */
    real (*iclt)[3][256] = malloc(sizeof(real) * ports * 3 * 256);
    real (*deb)[3][256] = malloc(sizeof(real) * ports * 3 * 256);
    for (int ip = 0; ip < ports; ip++) {
        for (int chn = 0; chn < numcol; chn++) {
            real clt[4][64];
            load_clt(&c->f, c->clt, ip, tile, chn, clt);
            if (extra != 0 && !forced)
                fract_shift(clt, (real)(extra * c->prm.port_offsets[ip][0] / c->prm.corr_magic_scale),
                            (real)(extra * c->prm.port_offsets[ip][1] / c->prm.corr_magic_scale));
            iclt_tile(clt, (real)0.25, c->lpf, iclt[ip][chn]);
            tile_debayer_shot_corr(chn != 2, iclt[ip][chn], c->window2, (real)c->prm.min_shot,
                                   (real)c->prm.scale_shot, deb[ip][chn]);
        }
    }
    combine_rgba(ports, port_mask, deb, iclt, c->window2, c->col_weights, &c->prm, c->planes, out, stats);
    free(iclt);
    free(deb);
}

/* ================================ port geometry (double) ================================ */
/* CorrVector angles of sensor i (GC:1217-1269): sensor 3 tilt/azimuth/zoom are minus the sum of
 * the other three; rolls add the factory roll in degrees. */
static void corr_angles(const clt_port_geometry* p, int i, double* az, double* tl, double* rl, double* zm) {
    const double* v = p->extrinsic;
    *tl = (i < 3) ? v[0 + i] : -(v[0] + v[1] + v[2]);
    *az = (i < 3) ? v[3 + i] : -(v[3] + v[4] + v[5]);
    *rl = v[6 + i] + p->roll_deg[i] * M_PI / 180.0;
    *zm = (i < 3) ? v[10 + i] : -(v[10] + v[11] + v[12]);
}

static void mat3_mul(const double* a, const double* b, double* c) {
    double t[9];
    for (int r = 0; r < 3; r++)
        for (int k = 0; k < 3; k++) t[r * 3 + k] = a[r * 3] * b[k] + a[r * 3 + 1] * b[3 + k] + a[r * 3 + 2] * b[6 + k];
    memcpy(c, t, sizeof(t));
}

/* GC:1043-1076 getRotMatrices (rots = a_r * (a_t * a_az)) and GC:1083-1147 getRotDeriveMatrices
 * (d/d azimuth, tilt, roll, zoom), ROT_AZ_SGN = -1, ROT_TL_SGN = ROT_RL_SGN = 1 (GC:1012-1014). */
static void rot_matrices(const clt_port_geometry* p, double rots[4][9], double drots[4][4][9]) {
    for (int i = 0; i < 4; i++) {
        double az, tl, rl, zm;
        corr_angles(p, i, &az, &tl, &rl, &zm);
        const double ca = cos(az), sa = sin(az), ct = cos(tl), st = sin(tl);
        const double zoom = 1.0 + zm, cr = cos(rl), sr = sin(rl);
        const double a_az[9] = {ca, 0, -sa, 0, 1, 0, sa, 0, ca};
        const double a_t[9] = {1, 0, 0, 0, ct, st, 0, -st, ct};
        const double a_r[9] = {cr * zoom, sr * zoom, 0, -sr * zoom, cr * zoom, 0, 0, 0, 1};
        const double a_daz[9] = {-sa, 0, -ca, 0, 0, 0, ca, 0, -sa};
        const double a_dt[9] = {0, 0, 0, 0, -st, ct, 0, -ct, -st};
        const double a_dr[9] = {-sr * zoom, cr * zoom, 0, -cr * zoom, -sr * zoom, 0, 0, 0, 0};
        const double a_dz[9] = {cr, sr, 0, -sr, cr, 0, 0, 0, 0};
        double t[9];
        mat3_mul(a_t, a_az, t);
        mat3_mul(a_r, t, rots[i]);
        if (!drots) continue;
        double u[9];
        mat3_mul(a_t, a_daz, u);
        mat3_mul(a_r, u, drots[i][0]);
        mat3_mul(a_dt, a_az, u);
        mat3_mul(a_r, u, drots[i][1]);
        mat3_mul(a_dr, t, drots[i][2]);
        mat3_mul(a_dz, t, drots[i][3]);
    }
}

/* GC:2707-2756 calcReverseDistortionTable: Newton solution of rD = r * k(r) per step 0.001 up to
 * 2.0 distortion radii; table of r / rD (NaN past a failure). */
#define RDIST_STEP 0.001
#define RDIST_LEN 2001
static void reverse_distortion_table(const clt_port_geometry* p, double* tab) {
    const double C = p->distortion[0], B = p->distortion[1], A = p->distortion[2], A5 = p->distortion[3];
    const double A6 = p->distortion[4], A7 = p->distortion[5], A8 = p->distortion[6];
    const int use8 = (A8 != 0.0) || (A7 != 0.0) || (A6 != 0.0);
    const double d = 1.0 - A8 - A7 - A6 - A5 - A - B - C;
    for (int j = 1; j < RDIST_LEN; j++) tab[j] = NAN;
    tab[0] = 1.0 / d;
    double rprev = 0.0, drdr = 1.0;
    for (int i = 1; i < RDIST_LEN; i++) {
        const double rdist = RDIST_STEP * i;
        double r = rprev + RDIST_STEP / drdr;
        for (int it = 0; it < 1000; it++) {
            double k;
            if (use8) {
                k = (((((((A8)*r + A7) * r + A6) * r + A5) * r + A) * r + B) * r + C) * r + d;
                drdr = (((((((8 * A8) * r + 7 * A7) * r + 6 * A6) * r + 5 * A5) * r + 4 * A) * r + 3 * B) * r + 2 * C) * r + d;
            } else {
                k = (((A5 * r + A) * r + B) * r + C) * r + d;
                drdr = (((5 * A5 * r + 4 * A) * r + 3 * B) * r + 2 * C) * r + d;
            }
            const double rd = r * k;
            if (drdr < 0.1) return;   /* too much distortion: the rest stays NaN */
            if (fabs(rd - rdist) < 1e-8) break;
            r += (rdist - rd) / drdr;
        }
        rprev = r;
        tab[i] = r / rdist;
    }
}

/* GC:2802-2830 getRByRDist: linear interpolation in the table, NaN outside. */
static double r_by_rdist(const double* tab, double rdist) {
    if (rdist < 0) return NAN;
    const int index = (int)floor(rdist / RDIST_STEP);
    if (index >= RDIST_LEN - 1) return NAN;
    return tab[index] + (tab[index + 1] - tab[index]) * (rdist / RDIST_STEP - index);
}

/* GC:2159-2314 getPortsCoordinatesAndDerivatives (main camera, no rig offsets) for tile centre
 * (px, py) and disparity: xy[4][2]; deriv (optional) [8][14]: columns 0..12 the reference's
 * pXYderiv, column 13 d/d disparity (not in the reference; from the same chain). */
static void port_coords(const clt_port_geometry* pm, const double* tab, const clt_port_geometry* p,
                        const double (*rxy)[2], const double rots[4][9], const double drots[4][4][9], double px,
                        double py, double disparity, double* xy, double* deriv) {
    /* the tile centre is undistorted by the main camera's model (gc_main, GC:2179-2184); rxy: the
     * port offsets of this camera (with rig offsets for the auxiliary one, GC:2177) */
    const double pXcd = px - 0.5 * pm->width, pYcd = py - 0.5 * pm->height;
    const double rD = sqrt(pXcd * pXcd + pYcd * pYcd) * 0.001 * pm->pixel_size;
    const double rND2R = r_by_rdist(tab, rD / pm->distortion_radius);
    const double pXc = pXcd * rND2R, pYc = pYcd * rND2R;
    const double fl_pix = p->focal_length / (0.001 * p->pixel_size);
    const double ri_scale = 0.001 * p->pixel_size / p->distortion_radius;
    if (deriv) memset(deriv, 0, sizeof(double) * 8 * 14);
    for (int i = 0; i < 4; i++) {
        const double vi[3] = {pXc - disparity * rxy[i][0], pYc - disparity * rxy[i][1], fl_pix};
        const double* R = rots[i];
        const double rvi[3] = {R[0] * vi[0] + R[1] * vi[1] + R[2] * vi[2], R[3] * vi[0] + R[4] * vi[1] + R[5] * vi[2],
                               R[6] * vi[0] + R[7] * vi[1] + R[8] * vi[2]};
        const double norm_z = fl_pix / rvi[2];
        const double pXci = rvi[0] * norm_z, pYci = rvi[1] * norm_z;
        const double rNDi = sqrt(pXci * pXci + pYci * pYci);
        const double ri = rNDi * ri_scale;
        double rD2rND = 1.0, rri = 1.0;
        for (int j = 0; j < 7; j++) {
            rri *= ri;
            rD2rND += p->distortion[j] * (rri - 1.0);
        }
        xy[2 * i] = pXci * rD2rND + p->pxy0[i][0];
        xy[2 * i + 1] = pYci * rD2rND + p->pxy0[i][1];
        if (!deriv) continue;
        double drD2rND_dri = 0.0;
        rri = 1.0;
        for (int j = 0; j < 7; j++) {
            drD2rND_dri += p->distortion[j] * (j + 1) * rri;
            rri *= ri;
        }
        /* d(pXci, pYci) by a change dv of rvi: (dv0 * norm_z - pXci * dv2 / rvi2, ...) */
        double dx[5], dy[5];   /* by azimuth, tilt, roll, zoom, disparity */
        for (int a = 0; a < 5; a++) {
            double dv[3];
            if (a < 4) {
                const double* D = drots[i][a];
                for (int r = 0; r < 3; r++) dv[r] = D[r * 3] * vi[0] + D[r * 3 + 1] * vi[1] + D[r * 3 + 2] * vi[2];
            } else {   /* d vi / d disparity = (-rx, -ry, 0) */
                for (int r = 0; r < 3; r++) dv[r] = -(R[r * 3] * rxy[i][0] + R[r * 3 + 1] * rxy[i][1]);
            }
            const double dxc = dv[0] * norm_z - pXci * dv[2] / rvi[2];
            const double dyc = dv[1] * norm_z - pYci * dv[2] / rvi[2];
            /* the reference drops the radius change for roll (GC:2257-2260): it is 0 */
            const double dri = (a == 2) ? 0.0 : ri_scale / rNDi * (pXci * dxc + pYci * dyc);
            const double dk = drD2rND_dri * dri;
            dx[a] = dxc * rD2rND + pXci * dk;
            dy[a] = dyc * rD2rND + pYci * dk;
        }
        double* rx = deriv + (2 * i) * 14;
        double* ry = deriv + (2 * i + 1) * 14;
        if (i < 3) {   /* GC:2285-2302: sensor 3 depends on minus the sum of the others */
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
}

/* ================================ exported API ================================ */
EXPORT int orc_real_size(void) { return (int)sizeof(real); }

EXPORT void orc_rot_matrices(const clt_port_geometry* p, double* rots, double* drots) {
    double r[4][9], d[4][4][9];
    rot_matrices(p, r, d);
    memcpy(rots, r, sizeof(r));
    if (drots) memcpy(drots, d, sizeof(d));
}

EXPORT double orc_r_by_rdist(const clt_port_geometry* p, double rdist) {
    double tab[RDIST_LEN];
    reverse_distortion_table(p, tab);
    return r_by_rdist(tab, rdist);
}

/* n tiles: px, py, disparity [n]; xy [n][8]; deriv [n][8][14] or NULL */
EXPORT void orc_port_coords(const clt_port_geometry* p, int n, const double* px, const double* py,
                            const double* disparity, double* xy, double* deriv) {
    double r[4][9], d[4][4][9];
    double* tab = malloc(sizeof(double) * RDIST_LEN);
    rot_matrices(p, r, d);
    reverse_distortion_table(p, tab);
    for (int t = 0; t < n; t++)
        port_coords(p, tab, p, (const double(*)[2])p->rxy, (const double(*)[9])r, (const double(*)[4][9])d, px[t], py[t],
                    disparity[t], xy + 8 * t, deriv ? deriv + 8 * 14 * t : NULL);
    free(tab);
}

/* GC:770-803 RigOffset.getRotMatrix: roll (with zoom) * tilt * azimuth of the auxiliary camera */
static void rig_matrix(const clt_rig_offset* ro, double* m) {
    const double ca = cos(ro->aux_azimuth), sa = sin(ro->aux_azimuth);
    const double ct = cos(ro->aux_tilt), st = sin(ro->aux_tilt);
    const double zoom = 1.0 + ro->aux_zoom, cr = cos(ro->aux_roll) * zoom, sr = sin(ro->aux_roll) * zoom;
    const double a_az[9] = {ca, 0, -sa, 0, 1, 0, sa, 0, ca};   /* ROT_AZ_SGN = -1 */
    const double a_t[9] = {1, 0, 0, 0, ct, st, 0, -st, ct};
    const double a_r[9] = {cr, sr, 0, -sr, cr, 0, 0, 0, 1};
    double t[9];
    mat3_mul(a_t, a_az, t);
    mat3_mul(a_r, t, m);
}

/* GC:728-745 RigOffset.recalcRXY: the auxiliary ports per nominal disparity pixel of the auxiliary
 * camera: its own rXY rolled by aux_roll, plus the baseline vector over its disparity radius */
static void rig_rxy(const clt_port_geometry* aux, const clt_rig_offset* ro, double rxy[4][2]) {
    const double xc = ro->baseline * cos(ro->aux_angle) / ro->aux_disparity_radius;
    const double yc = ro->baseline * sin(ro->aux_angle) / ro->aux_disparity_radius;
    const double c = cos(ro->aux_roll), sn = sin(ro->aux_roll);
    for (int i = 0; i < 4; i++) {
        rxy[i][0] = xc + c * aux->rxy[i][0] + sn * aux->rxy[i][1];
        rxy[i][1] = yc - sn * aux->rxy[i][0] + c * aux->rxy[i][1];
    }
}

/* clt_bi_quad's port centres (ID:8970-8993): the main quad as getPortsCoordinatesAndDerivatives
 * of the main camera at the target disparity; the auxiliary quad with use_rig_offsets (GC:2159-2230
 * with the main camera's undistortion, rXY_aux, CorrVector rotations times the rig matrix,
 * getRotMatrices(rigMatrix) GC:1016-1026) at disparity * R_aux / R_main.  xy: [n][16]. */
EXPORT void orc_port_coords_rig(const clt_port_geometry* main, const clt_port_geometry* aux, const clt_rig_offset* ro,
                                int n, const double* px, const double* py, const double* disparity, double* xy) {
    double rm[4][9], ra[4][9], rig[9], rxy_aux[4][2];
    double* tab = malloc(sizeof(double) * RDIST_LEN);
    rot_matrices(main, rm, NULL);
    rot_matrices(aux, ra, NULL);
    rig_matrix(ro, rig);
    for (int i = 0; i < 4; i++) mat3_mul(ra[i], rig, ra[i]);
    rig_rxy(aux, ro, rxy_aux);
    reverse_distortion_table(main, tab);
    const double dscale = ro->aux_disparity_radius / ro->main_disparity_radius;
    for (int t = 0; t < n; t++) {
        port_coords(main, tab, main, (const double(*)[2])main->rxy, (const double(*)[9])rm, NULL, px[t], py[t],
                    disparity[t], xy + 16 * t, NULL);
        port_coords(main, tab, aux, (const double(*)[2])rxy_aux, (const double(*)[9])ra, NULL, px[t], py[t],
                    disparity[t] * dscale, xy + 16 * t + 8, NULL);
    }
    free(tab);
}

EXPORT void orc_hwindow(double* out8) {
    tables();
    for (int i = 0; i < 8; i++) out8[i] = HW[i];
}

EXPORT void orc_tables(int* fold_index, int* unfold_index, double* unfold_k) {
    tables();
    if (fold_index) memcpy(fold_index, FOLD_INDEX, sizeof(FOLD_INDEX));
    if (unfold_index) memcpy(unfold_index, UNFOLD_INDEX, sizeof(UNFOLD_INDEX));
    if (unfold_k) memcpy(unfold_k, UNFOLD_K, sizeof(UNFOLD_K));
}

/* which: 0 dttt_iv, 1 dttt_iie, 2 dttt_iiie */
EXPORT void orc_dttt(int which, const double* x, int mode, double* y) {
    tables();
    real xi[64], yo[64];
    for (int i = 0; i < 64; i++) xi[i] = (real)x[i];
    if (which == 0) dttt_iv(xi, mode, yo);
    else if (which == 1) dttt_iie(xi, mode, yo);
    else dttt_iiie(xi, mode, yo);
    for (int i = 0; i < 64; i++) y[i] = yo[i];
}

EXPORT void orc_fold_tile(const double* x256, double shift_h, double shift_v, int mode, double* y64) {
    tables();
    real fk[4][64][4], xi[256], yo[64];
    for (int i = 0; i < 256; i++) xi[i] = (real)x256[i];
    shifted_fold_coeffs((real)shift_h, (real)shift_v, fk);
    fold_tile(xi, mode, fk, yo);
    for (int i = 0; i < 64; i++) y64[i] = yo[i];
}

EXPORT void orc_unfold_tile(const double* x64, int mode, double* y256) {
    tables();
    real xi[64], yo[256];
    for (int i = 0; i < 64; i++) xi[i] = (real)x64[i];
    unfold_tile(xi, mode, yo);
    for (int i = 0; i < 256; i++) y256[i] = yo[i];
}

EXPORT void orc_lpf(double sigma, double* out64) {
    tables();
    real f[64];
    make_lpf(sigma, f);
    for (int i = 0; i < 64; i++) out64[i] = f[i];
}

EXPORT void orc_fract_shift(double* clt256, double sx, double sy) {
    real c[4][64];
    for (int i = 0; i < 256; i++) c[i >> 6][i & 63] = (real)clt256[i];
    fract_shift(c, (real)sx, (real)sy);
    for (int i = 0; i < 256; i++) clt256[i] = c[i >> 6][i & 63];
}

EXPORT void orc_convolve_tile(double* data256, const double* kern256) {
    real d[4][64], k[4][64];
    for (int i = 0; i < 256; i++) {
        d[i >> 6][i & 63] = (real)data256[i];
        k[i >> 6][i & 63] = (real)kern256[i];
    }
    convolve_tile(d, k);
    for (int i = 0; i < 256; i++) data256[i] = d[i >> 6][i & 63];
}

/* One camera/colour tile, CPU order [4][h*8+v], residual shift; fract_shift applied if do_shift */
EXPORT void orc_convert_one(const clt_geom* g, const float* image, int stride, const float* kernels,
                            const float* offsets, int color, double cx, double cy, int do_shift,
                            double* clt256, double* res2) {
    tables();
    real clt[4][64];
    convert_one(g, image, stride, kernels, offsets, color, cx, cy, clt, res2);
    if (do_shift) fract_shift(clt, (real)res2[0], (real)res2[1]);
    for (int i = 0; i < 256; i++) clt256[i] = clt[i >> 6][i & 63];
}

EXPORT void orc_iclt_tile(const double* clt256, const double* lpf64, double* out256) {
    tables();
    real c[4][64], lpf[64], o[256];
    for (int i = 0; i < 256; i++) c[i >> 6][i & 63] = (real)clt256[i];
    if (lpf64)
        for (int i = 0; i < 64; i++) lpf[i] = (real)lpf64[i];
    iclt_tile(c, (real)0.25, lpf64 ? lpf : NULL, o);
    for (int i = 0; i < 256; i++) out256[i] = o[i];
}

EXPORT void orc_corr_unfold(const double* q256, double* out225) {
    real q[4][64], r[225];
    for (int i = 0; i < 256; i++) q[i >> 6][i & 63] = (real)q256[i];
    corr_unfold_tile(q, r);
    for (int i = 0; i < 225; i++) out225[i] = r[i];
}

EXPORT void orc_strip(const double* corr225, int hwidth, int dir, int ss, double* strip) {
    real c[225], s[16 * 15];
    for (int i = 0; i < 225; i++) c[i] = (real)corr225[i];
    strip_single(c, hwidth, dir, ss, s);
    for (int i = 0; i < hwidth * 15; i++) strip[i] = s[i];
}

EXPORT int orc_combine_strips(const double* strips, int nstrips, int len, double offset, int twice_diagonal,
                              double* combo) {
    real* sp[CLT_MAX_PAIRS];
    real* buf = malloc(sizeof(real) * (size_t)nstrips * len);
    real* cb = malloc(sizeof(real) * (size_t)len);
    for (int p = 0; p < nstrips; p++) {
        sp[p] = buf + (size_t)p * len;
        for (int i = 0; i < len; i++) sp[p][i] = (real)strips[(size_t)p * len + i];
    }
    int r = combine_strips(sp, nstrips, len, (real)offset, twice_diagonal, cb);
    for (int i = 0; i < len; i++) combo[i] = cb[i];
    free(buf);
    free(cb);
    return r;
}

EXPORT int orc_max_xy_int_axis(const double* data15, double min_max, int* ix) {
    real d[15];
    for (int i = 0; i < 15; i++) d[i] = (real)data15[i];
    return max_xy_int_axis(d, (real)min_max, ix);
}

EXPORT void orc_half_flat_top_window(int ihwidth, double hwidth, double blur, int normalize, int notch,
                                     double scale, double* wnd) {
    half_flat_top_window(ihwidth, hwidth, blur, normalize, notch, scale, wnd);
}

EXPORT int orc_max_x_cm(const double* strip, int height, int ixcenter, double* rslt3) {
    double wy[2], wx[9];
    half_flat_top_window(2, 1.0, 1.0, 1, 0, 1.0, wy);   /* IDP:77-79 corr_wndy_* */
    half_flat_top_window(9, 6.0, 5.0, 1, 0, 2.0, wx);   /* IDP:80-82 corr_wndx_*, wndx_scale 2 */
    real s[16 * 15], r[3];
    for (int i = 0; i < height * 15; i++) s[i] = (real)strip[i];
    int ok = max_x_cm(s, height, ixcenter, wy, 2, wx, 9, r);
    for (int i = 0; i < 3; i++) rslt3[i] = ok ? r[i] : NAN;
    return ok;
}

EXPORT int orc_max_xs_ortho(const double* corrs, int npairs, double offset, int is_vert, double enh_w,
                            double enh_s, double* rslt2) {
    double notch[15];
    ortho_notch(enh_w, enh_s, notch);
    real* sel[CLT_MAX_PAIRS];
    real* buf = malloc(sizeof(real) * 225 * (size_t)(npairs > 0 ? npairs : 1));
    for (int p = 0; p < npairs; p++) {
        sel[p] = buf + p * 225;
        for (int i = 0; i < 225; i++) sel[p][i] = (real)corrs[p * 225 + i];
    }
    real r[2];
    int ok = max_xs_ortho(sel, npairs, (real)offset, is_vert, notch, r);
    rslt2[0] = ok ? r[0] : NAN;
    rslt2[1] = ok ? r[1] : NAN;
    free(buf);
    return ok;
}

EXPORT void orc_tile_debayer(int rb, const double* tile, double* out) {
    real t[256], o[256];
    for (int i = 0; i < 256; i++) t[i] = (real)tile[i];
    tile_debayer(rb, t, o);
    for (int i = 0; i < 256; i++) out[i] = o[i];
}

/* ---- frame level ---- */
/* tile_combine_rgba + the final mix (combine_rgba) of one tile: deb, iclt [ports][3][256],
 * lt_window [256], chw [3] colour weights; out [planes][256], stats [4 * ports] or NULL. */
EXPORT void orc_tile_combine_rgba(int ports, int port_mask, const double* deb, const double* iclt, const double* lt_window,
                                  const double* chw, const clt_tex_params* prm, int planes, double* out, double* stats) {
    real (*d)[3][256] = malloc(sizeof(real) * ports * 3 * 256);
    real (*ic)[3][256] = malloc(sizeof(real) * ports * 3 * 256);
    real w[256], cw[3];
    for (int i = 0; i < ports * 3 * 256; i++) {
        ((real*)d)[i] = (real)deb[i];
        ((real*)ic)[i] = (real)iclt[i];
    }
    for (int i = 0; i < 256; i++) w[i] = (real)lt_window[i];
    for (int i = 0; i < 3; i++) cw[i] = (real)chw[i];
    combine_rgba(ports, port_mask, d, ic, w, cw, prm, planes, out, stats);
    free(d);
    free(ic);
}

/* dust_remove (ID:3002-3019) in place on pw [np][256]. */
EXPORT void orc_dust_remove(double* pw, int np) { dust_remove((wreal(*)[256])pw, np); }

EXPORT int orc_convert_frame(const clt_geom* g, const float* const* images, int stride, const float* const* kernels,
                             const float* const* offsets, const int32_t* tasks, int ntasks, int lpf_mask,
                             const double* lpf3x64, int no_fract_shift, double* clt_out, double* residual_out,
                             int nthreads) {
    tables();
    convert_ctx c;
    frame_init(&c.f, g);
    c.images = images;
    c.stride = stride;
    c.kernels = kernels;
    c.offsets = offsets;
    c.tasks = tasks;
    c.lpf_mask = lpf_mask;
    c.lpf = lpf3x64;
    c.no_fract_shift = no_fract_shift;
    c.clt = clt_out;
    c.residual = residual_out;
    run_pool(ntasks, nthreads, convert_task, &c);
    return 0;
}

/* valid: [tilesY*tilesX] bytes; out: [cams][colors][(H+8)*(W+8)] accumulated (must be zeroed) */
EXPORT int orc_imclt_frame(const clt_geom* g, const double* clt, const uint8_t* valid, const double* lpf64,
                           double* out, int nthreads) {
    tables();
    frame_t f;
    frame_init(&f, g);
    real lpf[64];
    if (lpf64)
        for (int i = 0; i < 64; i++) lpf[i] = (real)lpf64[i];
    size_t plane = (size_t)(g->height + 8) * (g->width + 8);
    real* buf = malloc(sizeof(real) * plane);
    for (int cam = 0; cam < g->num_cams; cam++)
        for (int col = 0; col < g->num_colors; col++) {
            for (size_t i = 0; i < plane; i++) buf[i] = 0;
            for (int pass = 0; pass < 4; pass++) {   /* ID:3588-3664: 4-pass checkerboard */
                imclt_ctx c = {f, clt, valid, lpf64 ? lpf : NULL, cam, col, pass, buf};
                int nx = (f.tiles_x + 1 - (pass & 1)) / 2;
                int ny = (f.tiles_y + 1 - ((pass >> 1) & 1)) / 2;
                run_pool(nx * ny, nthreads, imclt_tile_fn, &c);
            }
            double* o = out + ((size_t)cam * g->num_colors + col) * plane;
            for (size_t i = 0; i < plane; i++) o[i] = buf[i];
        }
    free(buf);
    return 0;
}

EXPORT int orc_corr_frame(const clt_geom* g, const double* clt, const int32_t* tasks, int ntasks,
                          const clt_pair* pairs, int npairs, const clt_corr_params* prm,
                          const clt_corr_lma_params* lma, double corr_sigma,
                          double* corr_out, double* combo_out, double* disp_out, int nthreads) {
    tables();
    if (npairs > CLT_MAX_PAIRS || prm->strip_height > 16) return CLT_E_ARG;
    if (lma && lma->enable && (lma->cnvx_hwnd_size < 1 || lma->cnvx_hwnd_size > LMA_HWIN_MAX)) return CLT_E_ARG;
    corr_ctx* c = calloc(1, sizeof(corr_ctx));
    frame_init(&c->f, g);
    c->clt = clt;
    c->tasks = tasks;
    c->pairs = pairs;
    c->npairs = npairs;
    c->prm = *prm;
    if (lma) c->lma = *lma;
    else c->lma.enable = 0;
    make_lpf(corr_sigma, c->lpf);
    color_weights(g, prm->corr_red, prm->corr_blue, c->col_weights);
    half_flat_top_window(2, 1.0, 1.0, 1, 0, 1.0, c->wndy);
    c->wndy_len = 2;
    half_flat_top_window(9, 6.0, 5.0, 1, 0, 2.0, c->wndx);
    c->wndx_len = 9;
    ortho_notch(prm->enhortho_width, prm->enhortho_scale, c->notch);
    c->corr_out = corr_out;
    c->combo_out = combo_out;
    c->disp_out = disp_out;
    run_pool(ntasks, nthreads, corr_task, c);
    free(c);
    return 0;
}

typedef struct {
    const clt_pair* pairs;
    int npairs;
    const clt_corr_lma_params* lp;
    const float* corrs;
    double* disp;
} lma_ctx;

static void lma_task(void* vctx, int t) {
    lma_ctx* c = (lma_ctx*)vctx;
    double (*dc)[225] = malloc(sizeof(double) * 225 * c->npairs);
    const float* src = c->corrs + (size_t)t * c->npairs * 225;
    for (int p = 0; p < c->npairs; p++)
        for (int i = 0; i < 225; i++) dc[p][i] = src[p * 225 + i];
    lma_chain((const double (*)[225])dc, c->pairs, c->npairs, c->lp, c->disp + (size_t)t * CLT_DISP_LAYERS);
    free(dc);
}

/* The LMA chain alone (ID:2254-2336) on given pair correlations [T][npairs][225] and disparity
 * records [T][CLT_DISP_LAYERS] (reads CM_RAW / STR_RAW, rewrites CM, STRENGTH, POLY, POLY_STR). */
EXPORT int orc_corr_lma_frame(const clt_pair* pairs, int npairs, const clt_corr_lma_params* lp, const float* corrs,
                              int ntasks, double* disp, int nthreads) {
    if (npairs > CLT_MAX_PAIRS || lp->cnvx_hwnd_size < 1 || lp->cnvx_hwnd_size > LMA_HWIN_MAX) return CLT_E_ARG;
    lma_ctx c = {pairs, npairs, lp, corrs, disp};
    run_pool(ntasks, nthreads, lma_task, &c);
    return 0;
}

/* Single fits for the oracle checks: corr_lma over the pairs flagged in sel (C2:1754-1900),
 * out = {disparity, strength, Wm, Wxy}; filter_convex (C2:2029-2207). */
EXPORT int orc_corr_lma(const double* corrs, const clt_pair* pairs, int npairs, const char* sel, double xcenter,
                        const clt_corr_lma_params* lp, double* out) {
    if (npairs > CLT_MAX_PAIRS || lp->cnvx_hwnd_size < 1 || lp->cnvx_hwnd_size > LMA_HWIN_MAX) return CLT_E_ARG;
    return corr_lma((const double (*)[225])corrs, pairs, npairs, sel, xcenter, lp, out);
}

EXPORT void orc_filter_convex(const double* d, int hwin, int x0c, int y0c, int add3x3, double nc_cost, double* w) {
    filter_convex(d, hwin, x0c, y0c, add3x3, nc_cost, w);
}

EXPORT double orc_foreground_correct(const clt_corr_lma_params* lp, const int* have, const double* dds, double full) {
    double d[4][4];
    memcpy(d, dds, sizeof(d));
    return foreground_correct(lp, have, d, full);
}

EXPORT int orc_textures_frame(const clt_geom* g, const double* clt, const int32_t* tasks, int ntasks,
                              const clt_tex_params* prm, const clt_corr_params* cprm, double corr_sigma,
                              const double* extra_disparity, double* tiles_out, double* stats_out, int nthreads) {
    tables();
    if (g->num_colors != 3) return CLT_E_UNSUPPORTED;
    tex_ctx* c = calloc(1, sizeof(tex_ctx));
    frame_init(&c->f, g);
    c->clt = clt;
    c->tasks = tasks;
    c->extra_disparity = extra_disparity;
    c->prm = *prm;
    make_lpf(corr_sigma, c->lpf);
    color_weights(g, cprm->corr_red, cprm->corr_blue, c->col_weights);
    for (int i = 0; i < 16; i++)   /* DR:695-712 getWin2d, squared (ID:1725-1726) */
        for (int j = 0; j < 16; j++) {
            double w = HW[(i < 8) ? i : 15 - i] * HW[(j < 8) ? j : 15 - j];
            c->window2[i * 16 + j] = (real)(w * w);
        }
    c->planes = prm->keep_weights ? 4 + g->num_cams + 4 : 4;
    c->tiles_out = tiles_out;
    c->stats_out = stats_out;
    run_pool(ntasks, nthreads, tex_task, c);
    free(c);
    return 0;
}

/* ID:3855-3991 combineRGBATiles on the texture tiles of a task list: the tile of a position is the
 * last task on it with a non-zero task word (texture_tiles[tileY][tileX]).  overlap: the 16x16 tiles
 * at stride 8, the tile's pixel (i, j) added at (8 ty + i - 4, 8 tx + j - 4) inside the frame; with
 * sharp_alpha the alpha plane (3) takes only the centre 8x8 of each tile (ID:3931-3937).  Without
 * overlap: each tile copied as a 16x16 block at (16 ty, 16 tx) of a (2 height) x (2 width) mosaic.
 * out: [planes][H][W] (overlap) or [planes][2H][2W], zeroed by the caller. */
EXPORT int orc_combine_textures(const clt_geom* g, const int32_t* tasks, int ntasks, const double* tiles,
                                int planes, int overlap, int sharp_alpha, double* out) {
    frame_t f;
    frame_init(&f, g);
    const int W = g->width, H = g->height, TX = f.tiles_x, TY = f.tiles_y;
    int* map = malloc(sizeof(int) * (size_t)TX * TY);
    for (int i = 0; i < TX * TY; i++) map[i] = -1;
    for (int t = 0; t < ntasks; t++) {
        if (task_flags(g, tasks, t) == 0) continue;
        int tx, ty;
        task_txy(g, tasks, t, &tx, &ty);
        map[ty * TX + tx] = t;
    }
    const int n2 = 16, n_half = 4;
    for (int ty = 0; ty < TY; ty++)
        for (int tx = 0; tx < TX; tx++) {
            const int t = map[ty * TX + tx];
            if (t < 0) continue;
            const double* tt = tiles + (size_t)t * planes * 256;
            if (!overlap) {
                for (int i = 0; i < n2; i++)
                    for (int p = 0; p < planes; p++)
                        for (int j = 0; j < n2; j++)
                            out[((size_t)p * 2 * H + ty * n2 + i) * 2 * W + tx * n2 + j] = tt[p * 256 + i * n2 + j];
                continue;
            }
            for (int i = 0; i < n2; i++) {
                const int y = ty * 8 + i - n_half;
                if (y < 0 || y >= H) continue;
                for (int p = 0; p < planes; p++) {
                    const int sharp = (p == 3) && sharp_alpha;
                    if (sharp && (i < n_half || i >= n2 - n_half)) continue;
                    for (int j = sharp ? n_half : 0; j < (sharp ? n2 - n_half : n2); j++) {
                        const int x = tx * 8 + j - n_half;
                        if (x < 0 || x >= W) continue;
                        out[((size_t)p * H + y) * W + x] += tt[p * 256 + i * n2 + j];
                    }
                }
            }
        }
    free(map);
    return 0;
}

/* ================= two-quad rig (ImageDtt.clt_bi_quad, ID:8731-9504) ================= */

/* Jama 1.0.3 LUDecomposition (Matrix.det / Matrix.solve of PolynomialApproximation; Jama is not in
 * the reference) restated as textbook LU with partial pivoting.  Returns the determinant; b (may
 * be NULL) is replaced by A^-1 b when the determinant is not 0.  a (n x n, n <= 6) is destroyed. */
static double lu_det_solve(int n, double* a, double* b) {
    double det = 1.0;
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(a[i * n + k]) > fabs(a[p * n + k])) p = i;
        if (p != k) {
            for (int j = 0; j < n; j++) {
                double t = a[k * n + j];
                a[k * n + j] = a[p * n + j];
                a[p * n + j] = t;
            }
            if (b) {
                double t = b[k];
                b[k] = b[p];
                b[p] = t;
            }
            det = -det;
        }
        const double d = a[k * n + k];
        det *= d;
        if (d == 0.0) continue;
        for (int i = k + 1; i < n; i++) {
            const double f = a[i * n + k] / d;
            for (int j = k + 1; j < n; j++) a[i * n + j] -= f * a[k * n + j];
            if (b) b[i] -= f * b[k];
        }
    }
    if (b && det != 0.0)
        for (int k = n - 1; k >= 0; k--) {
            double x = b[k];
            for (int j = k + 1; j < n; j++) x -= a[k * n + j] * b[j];
            b[k] = x / a[k * n + k];
        }
    return det;
}

/* PA:590-598 normMatix: the product of the row norms */
static double norm_matrix(int n, const double* a) {
    double norm = 1.0;
    for (int i = 0; i < n; i++) {
        double d = 0.0;
        for (int j = 0; j < n; j++) d += a[i * n + j] * a[i * n + j];
        norm *= sqrt(d);
    }
    return norm;
}

/* PA:201-224 quadraticMaxV2dX2Y2XY over PA:350-580 quadraticApproximation (forceLinear false,
 * no damping, thresholdLin 1e-20): weighted least squares z ~ A x^2 + B y^2 + C xy + D x + E y + F
 * over the samples with w > 0, then the stationary point.  Returns 0 where the reference returns
 * null; out = {x, y, f(x, y), A, B}. */
static int quadratic_max(const double* x, const double* y, const double* z, const double* w, int n,
                         double thr_quad, double out[5]) {
    if (n == 0) return 0;
    double S00 = 0, S10 = 0, S01 = 0, S20 = 0, S11 = 0, S02 = 0, S30 = 0, S21 = 0, S12 = 0, S03 = 0;
    double S40 = 0, S31 = 0, S22 = 0, S13 = 0, S04 = 0, Z00 = 0, Z10 = 0, Z01 = 0, Z20 = 0, Z11 = 0, Z02 = 0;
    for (int i = 0; i < n; i++) {
        const double wi = w[i];
        if (!(wi > 0)) continue;
        const double xi = x[i], yi = y[i], x2 = xi * xi, y2 = yi * yi;
        S00 += wi;
        S10 += wi * xi;
        S01 += wi * yi;
        S11 += wi * xi * yi;
        S20 += wi * x2;
        S02 += wi * y2;
        S30 += wi * x2 * xi;
        S21 += wi * x2 * yi;
        S12 += wi * xi * y2;
        S03 += wi * y2 * yi;
        S40 += wi * x2 * x2;
        S31 += wi * x2 * xi * yi;
        S22 += wi * x2 * y2;
        S13 += wi * xi * y2 * yi;
        S04 += wi * y2 * y2;
        const double wz = wi * z[i];
        Z00 += wz;
        Z10 += wz * xi;
        Z01 += wz * yi;
        Z20 += wz * x2;
        Z11 += wz * xi * yi;
        Z02 += wz * y2;
    }
    double ml[9] = {S20, S11, S10, S11, S02, S01, S10, S01, S00};
    const double nml = norm_matrix(3, ml);
    if (nml == 0.0 || fabs(lu_det_solve(3, ml, NULL)) / nml < 1e-20) return 0;   /* linear only: null */
    double mq[36] = {S40, S22, S31, S30, S21, S20, S22, S04, S13, S12, S03, S02, S31, S13, S22, S21, S12, S11,
                     S30, S12, S21, S20, S11, S10, S21, S03, S12, S11, S02, S01, S20, S02, S11, S10, S01, S00};
    const double nmq = norm_matrix(6, mq);
    double c[6] = {Z20, Z02, Z11, Z10, Z01, Z00};
    if (nmq == 0.0 || fabs(lu_det_solve(6, mq, c)) / nmq < thr_quad) return 0;
    double m[4] = {2 * c[0], c[2], c[2], 2 * c[1]};
    const double nm = norm_matrix(2, m);
    double xy[2] = {-c[3], -c[4]};
    if (nm == 0.0 || fabs(lu_det_solve(2, m, xy)) / nm < thr_quad) return 0;
    out[0] = xy[0];
    out[1] = xy[1];
    out[2] = c[0] * xy[0] * xy[0] + c[1] * xy[1] * xy[1] + c[2] * xy[0] * xy[1] + c[3] * xy[0] + c[4] * xy[1] + c[5];
    out[3] = c[0];
    out[4] = c[1];
    return 1;
}

/* C2:1509-1536 single2dPoly: corrLMA(run_poly_instead = true) of one correlation as a horizontal
 * pair (C2:1903-1935), getMaxXYPoly (Correlations2dLMA:643-705) over its samples; out = {xcenter -
 * x, -y, value} or {NaN, NaN, 0}. */
static void single2d_poly(const double* corr, double xcenter, const clt_corr_lma_params* lp, double out[3]) {
    static const clt_pair hor = {0, 1, 0, 1};
    const char sel = 1;
    out[0] = out[1] = NAN;
    out[2] = 0.0;
    lma_t* L = malloc(sizeof(lma_t));
    lma_samples(L, (const double (*)[225])corr, &hor, 1, &sel, xcenter, lp);
    double q[5];
    if (quadratic_max(L->x, L->y, L->v, L->w, L->ns, 1.0e-30, q) && !(q[2] < 0.0) && !(q[3] >= 0.0) && !(q[4] >= 0.0)) {
        out[0] = xcenter - q[0];
        out[1] = -q[1];
        out[2] = isnan(q[2]) ? 0.0 : q[2];
    }
    free(L);
}

/* C2:1538-1586 single2dCM: centre of mass of the positive values within radius of (round(xcenter), 0);
 * out = {xcenter - xm, -ym, value at the rounded centre} */
static void single2d_cm(const double* corr, double xcenter, double radius, double out[3]) {
    const int width = 15, center = 7, center_index = (width + 1) * center;
    const int ixcenter = (int)jround(xcenter);
    const int iradius = (int)fabs(radius), ir2 = (int)(radius * radius), square = radius < 0;
    double s0 = 0, sx = 0, sy = 0;
    for (int y = -iradius; y <= iradius; y++) {
        const int dy = y;
        if (dy < -center || dy > center) continue;
        for (int x = -iradius; x <= iradius; x++) {
            const int dx = ixcenter + x;
            const double r2 = y * y + x * x;
            if (dx >= -center && dx <= center && (square || r2 <= ir2)) {
                const double d = corr[dy * width + dx + center_index];
                if (d > 0.0) {
                    s0 += d;
                    sx += d * dx;
                    sy += d * dy;
                }
            }
        }
    }
    const double xm = sx / s0, ym = sy / s0;
    const int ixm = isnan(xm) ? 0 : (int)jround(xm), iym = isnan(ym) ? 0 : (int)jround(ym);   /* Math.round(NaN) = 0 */
    out[0] = xcenter - xm;
    out[1] = -ym;
    out[2] = corr[iym * width + ixm + center_index];
}

typedef struct {
    frame_t f;
    const int32_t* tasks;
    int ntasks;
    clt_corr_params prm;
    clt_corr_lma_params lp;   /* mix_corr_poly = use_poly, no foreground correction */
    clt_rig_params rig;
    double wndy[16], wndx[16], wndy_notch[32];
    const double *corr_main, *disp_main, *corr_aux, *disp_aux, *inter;
    const float* target;
    const int32_t* tile_task;
    double* bimap;
} rig_ctx;

/* ID:7370-7560 tileInterCamCorrs on one inter-camera correlation: the single correlation as a
 * horizontal pair's strip, the integer maximum, the centre of mass (or its notch form), the LMA
 * (C2:1903-1935) and dx, dy; res = {disparity, strength, dx, dy} */
static void inter_cam_corrs(const rig_ctx* c, const double* corr, int no_int_x0, double res[4]) {
    const int hw = c->prm.strip_height;
    res[0] = NAN;
    res[1] = 0.0;
    res[2] = res[3] = NAN;
    real rc[225], strip[16 * 15], combo[16 * 15];
    for (int i = 0; i < 225; i++) rc[i] = (real)corr[i];
    strip_single(rc, hw, 0, 1, strip);
    real* strips[1] = {strip};
    combine_strips(strips, 1, hw * 15, (real)c->prm.corr_offset, c->prm.twice_diagonal, combo);
    int ix;
    if (!max_xy_int_axis(combo, (real)((no_int_x0 ? 0.5 : 1.0) * c->prm.min_corr), &ix)) return;
    res[0] = -ix;
    res[1] = combo[ix + 7];
    real cs[3];
    const int notch = c->rig.notch_mode != 0;
    const int ok = notch ? max_x_cm(combo, hw, 0, c->wndy_notch, c->rig.corr_strip_notch, c->wndx, 9, cs)
                         : max_x_cm(combo, hw, no_int_x0 ? 0 : ix, c->wndy, 2, c->wndx, 9, cs);
    if (!ok) return;
    res[0] = -cs[0];
    const double eff_radius = cs[2] * c->rig.cm_max_normalization;
    const double strength = cs[1] / (eff_radius * eff_radius);
    res[1] = strength;
    if (!(strength > c->lp.min_poly_strength) || notch) return;
    static const clt_pair hor = {0, 1, 0, 1};
    const char sel = 1;
    double r[4];
    if (!corr_lma((const double (*)[225])corr, &hor, 1, &sel, cs[0], &c->lp, r)) return;
    if (c->rig.use_poly) {
        res[0] = r[0];
        res[1] = r[1];
    }
    double xy[3];
    if (c->rig.use_xy_poly) single2d_poly(corr, 0.0, &c->lp, xy);
    else single2d_cm(corr, 0.0, c->rig.max_corr_radius, xy);
    res[2] = xy[0];
    res[3] = xy[1];
}

/* ID:7581-7740 tileCorrs of one quad from its pair correlations and its correlate2D record
 * (CM_RAW = the centre of mass, STR_RAW = the strength at the integer maximum): the LMA over all
 * pairs (replacing disparity/strength when use_poly) and the four one-direction LMA fits;
 * res = {disparity, hor, vert, diag main, diag other, strength, ... the same strengths} */
static void tile_corrs(const rig_ctx* c, const double (*corrs)[225], const double* rec, double res[10]) {
    static const clt_pair quad[6] = {{0, 1, 0, 1}, {2, 3, 0, 1}, {0, 2, 1, 1}, {1, 3, 1, 1}, {0, 3, 2, 1}, {2, 1, 3, 1}};
    for (int i = 0; i < 5; i++) res[i] = NAN;
    for (int i = 5; i < 10; i++) res[i] = 0.0;
    const double cm = rec[CLT_DISP_CM_RAW], str = rec[CLT_DISP_STR_RAW];
    res[0] = isnan(cm) ? rec[CLT_DISP_INT] : cm;   /* -ixy when the centre of mass has no weight */
    res[5] = str;
    if (isnan(cm) || !(str > c->lp.min_poly_strength)) return;
    char sel[6] = {1, 1, 1, 1, 1, 1};
    double r[4];
    if (!corr_lma(corrs, quad, 6, sel, -cm, &c->lp, r)) return;
    double disparity = cm;
    if (c->rig.use_poly) {
        disparity = r[0];
        res[0] = r[0];
        res[5] = r[1];
    }
    for (int dir = 0; dir < 4; dir++) {   /* C2:1591-1630 corr4dirsLMA */
        for (int p = 0; p < 6; p++) sel[p] = quad[p].dir == dir;
        double dd[4];
        if (corr_lma(corrs, quad, 6, sel, -disparity, &c->lp, dd)) {
            res[1 + dir] = dd[0];
            res[6 + dir] = dd[1];
        }
    }
}

static const int SNGL_TO_BI[2][10] = {   /* ID:190-194 */
    {CLT_BI_DISP_FULL, CLT_BI_DISP_HOR, CLT_BI_DISP_VERT, CLT_BI_DISP_DIAGM, CLT_BI_DISP_DIAGO, CLT_BI_STR_FULL,
     CLT_BI_STR_HOR, CLT_BI_STR_VERT, CLT_BI_STR_DIAGM, CLT_BI_STR_DIAGO},
    {CLT_BI_ADISP_FULL, CLT_BI_ADISP_HOR, CLT_BI_ADISP_VERT, CLT_BI_ADISP_DIAGM, CLT_BI_ADISP_DIAGO, CLT_BI_ASTR_FULL,
     CLT_BI_ASTR_HOR, CLT_BI_ASTR_VERT, CLT_BI_ASTR_DIAGM, CLT_BI_ASTR_DIAGO}};

static void put_inter(double* b, const double r[4]) {   /* ID:9238-9248 */
    b[CLT_BI_DISP_CROSS] = r[0];
    b[CLT_BI_STR_CROSS] = r[1];
    b[CLT_BI_DISP_CROSS_DX] = r[2];
    b[CLT_BI_DISP_CROSS_DY] = r[3];
    b[CLT_BI_STR_ALL] = pow(r[1] * b[CLT_BI_STR_FULL] * b[CLT_BI_ASTR_FULL], 1.0 / 3);
}

/* ID:8946-9248 for one task: both quads' tileCorrs, the inter-camera result, the record */
static void rig_task(void* vctx, int t) {
    rig_ctx* c = (rig_ctx*)vctx;
    double* b = c->bimap + (size_t)t * CLT_BI_LAYERS;
    for (int i = 0; i < CLT_BI_LAYERS; i++) b[i] = 0.0;   /* the reference's zeroed arrays */
    if (task_flags(c->f.g, c->tasks, t) == 0) {
        b[CLT_BI_TARGET] = NAN;
        return;
    }
    b[CLT_BI_TARGET] = c->target ? c->target[t] : NAN;
    for (int q = 0; q < 2; q++) {
        const double* corrs = (q ? c->corr_aux : c->corr_main) + (size_t)t * 6 * 225;
        const double* rec = (q ? c->disp_aux : c->disp_main) + (size_t)t * CLT_DISP_LAYERS;
        double res[10];
        tile_corrs(c, (const double (*)[225])corrs, rec, res);
        for (int i = 0; i < 10; i++) b[SNGL_TO_BI[q][i]] = res[i];
    }
    double r[4];
    inter_cam_corrs(c, c->inter + (size_t)t * 225, c->rig.no_int_x0, r);
    put_inter(b, r);
}

/* ID:9404-9480: the inter-camera correlations averaged over the (2 lt_rad + 1)^2 neighbours
 * (vertical neighbours only in notch mode) with cos weights, then tileInterCamCorrs (no_int_x0) */
static void rig_lt_task(void* vctx, int t) {
    rig_ctx* c = (rig_ctx*)vctx;
    if (task_flags(c->f.g, c->tasks, t) == 0) return;
    int tx, ty;
    task_txy(c->f.g, c->tasks, t, &tx, &ty);
    const int lr = c->rig.lt_rad, lrx = c->rig.notch_mode ? 0 : lr;
    double avg[225], sw = 0.0;
    for (int i = 0; i < 225; i++) avg[i] = 0.0;
    for (int dy = -lr; dy <= lr; dy++)
        for (int dx = -lrx; dx <= lrx; dx++) {
            const int x = tx + dx, y = ty + dy;
            if (x < 0 || y < 0 || x >= c->f.tiles_x || y >= c->f.tiles_y) continue;
            const int o = c->tile_task[y * c->f.tiles_x + x];
            if (o < 0) continue;
            const double nw = cos(M_PI * abs(dy) / (2 * lr + 1)) * cos(M_PI * abs(dx) / (2 * lr + 1));
            const double* oc = c->inter + (size_t)o * 225;
            for (int i = 0; i < 225; i++) avg[i] += nw * oc[i];
            sw += nw;
        }
    if (!(sw > 0.0)) return;
    const double s = 1.0 / sw;
    for (int i = 0; i < 225; i++) avg[i] *= s;
    double r[4];
    inter_cam_corrs(c, avg, 1, r);
    put_inter(c->bimap + (size_t)t * CLT_BI_LAYERS, r);
}

/* The rig analysis from the pair correlations and correlate2D records of both quads
 * ([T][6][225], [T][CLT_DISP_LAYERS]) and the inter-camera correlations [T][225] -> bimap
 * [T][CLT_BI_LAYERS].  g: the frame (8 cameras); target: [T] or NULL. */
EXPORT int orc_rig_analysis(const clt_geom* g, const int32_t* tasks, int ntasks, const clt_corr_params* prm,
                            const clt_corr_lma_params* lma, const clt_rig_params* rig, const double* corr_main,
                            const double* disp_main, const double* corr_aux, const double* disp_aux,
                            const double* inter, const float* target, double* bimap, int nthreads) {
    if (g->num_cams != 8 || prm->strip_height > 16 || rig->corr_strip_notch > 32 || rig->corr_strip_notch < 1 ||
        lma->cnvx_hwnd_size < 1 || lma->cnvx_hwnd_size > LMA_HWIN_MAX)
        return CLT_E_ARG;
    rig_ctx* c = calloc(1, sizeof(rig_ctx));
    frame_init(&c->f, g);
    c->tasks = tasks;
    c->ntasks = ntasks;
    c->prm = *prm;
    c->lp = *lma;
    c->lp.mix_corr_poly = rig->use_poly;
    c->lp.fo_correct = 0;
    c->rig = *rig;
    half_flat_top_window(2, 1.0, 1.0, 1, 0, 1.0, c->wndy);   /* C2:139-159 with IDP defaults */
    half_flat_top_window(9, 6.0, 5.0, 1, 0, 2.0, c->wndx);
    half_flat_top_window(rig->corr_strip_notch, rig->corr_notch_hwidth, rig->corr_notch_blur, 1, 1, 2.0, c->wndy_notch);
    c->corr_main = corr_main;
    c->disp_main = disp_main;
    c->corr_aux = corr_aux;
    c->disp_aux = disp_aux;
    c->inter = inter;
    c->target = target;
    c->bimap = bimap;
    run_pool(ntasks, nthreads, rig_task, c);
    if (rig->lt_rad > 0) {
        int32_t* tt = malloc(sizeof(int32_t) * c->f.num_tiles);
        for (int i = 0; i < c->f.num_tiles; i++) tt[i] = -1;
        for (int t = 0; t < ntasks; t++)
            if (task_flags(g, tasks, t) != 0) {
                int tx, ty;
                task_txy(g, tasks, t, &tx, &ty);
                tt[ty * c->f.tiles_x + tx] = t;
            }
        c->tile_task = tt;
        run_pool(ntasks, nthreads, rig_lt_task, c);
        free(tt);
    }
    free(c);
    return 0;
}

typedef struct {
    frame_t f;
    const double* clt;
    const int32_t* tasks;
    real lpf[64], col_weights[3];
    real fat_zero;
    double* inter;
} inter_ctx;

/* C2:401-448 correlateInterCamerasFD: the mean CLT of each quad (cltMixCameras), then
 * correlateCompositeFD of the two means */
static void inter_task(void* vctx, int t) {
    inter_ctx* c = (inter_ctx*)vctx;
    double* o = c->inter + (size_t)t * 225;
    if (task_flags(c->f.g, c->tasks, t) == 0) {
        for (int i = 0; i < 225; i++) o[i] = NAN;
        return;
    }
    int tx, ty;
    task_txy(c->f.g, c->tasks, t, &tx, &ty);
    const size_t tile = (size_t)ty * c->f.tiles_x + tx;
    const int K = c->f.g->num_colors;
    real mix[2][3][4][64], blk[4][64];
    for (int q = 0; q < 2; q++)
        for (int col = 0; col < K; col++) {
            double acc[4][64] = {{0}};
            for (int cam = 0; cam < 4; cam++) {
                load_clt(&c->f, c->clt, 4 * q + cam, tile, col, blk);
                for (int n = 0; n < 4; n++)
                    for (int i = 0; i < 64; i++) acc[n][i] += blk[n][i];
            }
            for (int n = 0; n < 4; n++)
                for (int i = 0; i < 64; i++) mix[q][col][n][i] = (real)(acc[n][i] * 0.25);
        }
    real out[225];
    correlate_composite_fd(K, mix[0], mix[1], c->lpf, c->col_weights, c->fat_zero, out);
    for (int i = 0; i < 225; i++) o[i] = out[i];
}

/* clt_bi_quad (ID:8731-9504) correlation part over a converted 8-camera frame: both quads'
 * correlate2D (cameras 0..3 and 4..7, the quad pair table), the inter-camera correlations and the
 * rig analysis.  Outputs may not be NULL: corr_main/aux [T][6][225], disp_main/aux
 * [T][CLT_DISP_LAYERS], inter [T][225], bimap [T][CLT_BI_LAYERS]. */
EXPORT int orc_rig_frame(const clt_geom* g, const double* clt, const int32_t* tasks, int ntasks,
                         const clt_corr_params* prm, const clt_corr_lma_params* lma, const clt_rig_params* rig,
                         double corr_sigma, const float* target, double* corr_main, double* disp_main,
                         double* corr_aux, double* disp_aux, double* inter, double* bimap, int nthreads) {
    tables();
    if (g->num_cams != 8 || g->num_colors != 3) return CLT_E_ARG;
    const clt_pair qp[2][6] = {{{0, 1, 0, 1}, {2, 3, 0, 1}, {0, 2, 1, 1}, {1, 3, 1, 1}, {0, 3, 2, 1}, {2, 1, 3, 1}},
                               {{4, 5, 0, 1}, {6, 7, 0, 1}, {4, 6, 1, 1}, {5, 7, 1, 1}, {4, 7, 2, 1}, {6, 5, 3, 1}}};
    int rc = orc_corr_frame(g, clt, tasks, ntasks, qp[0], 6, prm, NULL, corr_sigma, corr_main, NULL, disp_main, nthreads);
    if (!rc) rc = orc_corr_frame(g, clt, tasks, ntasks, qp[1], 6, prm, NULL, corr_sigma, corr_aux, NULL, disp_aux, nthreads);
    if (rc) return rc;
    inter_ctx* ic = calloc(1, sizeof(inter_ctx));
    frame_init(&ic->f, g);
    ic->clt = clt;
    ic->tasks = tasks;
    make_lpf(corr_sigma, ic->lpf);
    color_weights(g, prm->corr_red, prm->corr_blue, ic->col_weights);
    ic->fat_zero = (real)prm->fat_zero;
    ic->inter = inter;
    run_pool(ntasks, nthreads, inter_task, ic);
    free(ic);
    return orc_rig_analysis(g, tasks, ntasks, prm, lma, rig, corr_main, disp_main, corr_aux, disp_aux, inter, target,
                            bimap, nthreads);
}

/* Correlation-peak offsets of quad pairs (the clt_mismatch measurement, ID:2342-2420, by the centre
 * of mass of mismatchPairsCM, C2:1336-1430; the window is centred at the pair's integer maximum,
 * the first in row order, instead of at the residual disparity): corr [n][6][225], out [n][6][3] =
 * dx, dy of the centre of mass of the positive values within radius (image axes, pixels from the
 * centre; NaN where none is positive), the correlation at the maximum. */
EXPORT void orc_corr_peak_offsets(const double* corr, int n, double radius, double* out) {
    const int iradius = (int)fabs(radius), ir2 = (int)(radius * radius);
    for (int k = 0; k < 6 * n; k++) {
        const double* c = corr + (size_t)k * 225;
        int im = 0;
        for (int i = 1; i < 225; i++)
            if (c[i] > c[im]) im = i;
        const int iy = im / 15 - 7, ix = im % 15 - 7;
        double s0 = 0, sx = 0, sy = 0;
        for (int y = iy - iradius; y <= iy + iradius; y++) {
            if (y < -7 || y > 7) continue;
            for (int x = ix - iradius; x <= ix + iradius; x++) {
                const int r2 = (x - ix) * (x - ix) + (y - iy) * (y - iy);
                if (x < -7 || x > 7 || r2 > ir2) continue;
                const double d = c[(y + 7) * 15 + x + 7];
                if (d > 0.0) {
                    s0 += d;
                    sx += d * x;
                    sy += d * y;
                }
            }
        }
        out[3 * k] = s0 > 0.0 ? sx / s0 : NAN;
        out[3 * k + 1] = s0 > 0.0 ? sy / s0 : NAN;
        out[3 * k + 2] = c[im];
    }
}

/* single checks for the tests */
EXPORT int orc_quadratic_max(const double* x, const double* y, const double* z, const double* w, int n, double* out5) {
    return quadratic_max(x, y, z, w, n, 1.0e-30, out5);
}
EXPORT void orc_single2d_cm(const double* corr, double xcenter, double radius, double* out3) {
    single2d_cm(corr, xcenter, radius, out3);
}
EXPORT void orc_single2d_poly(const double* corr, double xcenter, const clt_corr_lma_params* lp, double* out3) {
    single2d_poly(corr, xcenter, lp, out3);
}

/* ================= CLT kernel preparation (QuadCLT.calculateCLTKernel, QC:433-668) ================= */
/* ID:4625-4672 clt_convert_double_kernel: centroid of the double-resolution kernel, 15x15 downsample
 * around it with the 3x3 weights {1, 1/2 (sides), 1/4 (corners)}, then data_x, data_y (the
 * chosen centre, half-pixel steps) and center_x, center_y (the exact centroid) */
static void convert_double_kernel(const double* src, int src_size, double* dst) {
    const int dtt_size = 8;
    const int indices[9] = {0, -src_size, -1, 1, src_size, -src_size - 1, -src_size + 1, src_size - 1, src_size + 1};
    const double weights[9] = {1.0, 0.5, 0.5, 0.5, 0.5, 0.25, 0.25, 0.25, 0.25};
    const int src_center = src_size / 2;
    double sx = 0.0, sy = 0.0, s = 0.0;
    int indx = 0;
    for (int i = -src_center; i < src_center; i++)
        for (int j = -src_center; j < src_center; j++) {
            double d = src[indx++];
            sx += j * d;
            sy += i * d;
            s += d;
        }
    int src_x = (int)jround(sx / s) + src_center;
    int src_y = (int)jround(sy / s) + src_center;
    if (src_x < 2 * dtt_size) src_x = 2 * dtt_size - 1;
    else if (src_x > (src_size - 2 * dtt_size)) src_x = src_size - 2 * dtt_size;
    if (src_y < 2 * dtt_size) src_y = 2 * dtt_size - 1;
    else if (src_y > (src_size - 2 * dtt_size)) src_y = src_size - 2 * dtt_size;
    indx = 0;
    for (int i = -dtt_size + 1; i < dtt_size; i++) {
        int src_i = (src_y + 2 * i) * src_size + src_x;
        for (int j = -dtt_size + 1; j < dtt_size; j++) {
            double d = 0.0;
            for (int k = 0; k < 9; k++) d += weights[k] * src[src_i + 2 * j + indices[k]];
            dst[indx++] = d;
        }
    }
    dst[indx++] = 0.5 * (src_x - src_center);
    dst[indx++] = 0.5 * (src_y - src_center);
    dst[indx++] = 0.5 * (sx / s);
    dst[indx++] = 0.5 * (sy / s);
}

/* ID:4674-4696 clt_normalize_kernel: sum(kernel * window[|i|][|j|]) -> 16 */
static void normalize_kernel(double* k, const double* window) {
    double s = 0.0;
    int indx = 0;
    for (int i = -7; i < 8; i++)
        for (int j = -7; j < 8; j++) s += k[indx++] * window[abs(i) * 8 + abs(j)];
    s = 1.0 / s;
    for (int i = 0; i < 225; i++) k[i] *= 16 * s;
}

/* ID:4698-4723 clt_symmetrize_kernel (the fourth quadrant takes the second's signs, as there) */
static void symmetrize_kernel(const double* k, double sym[4][64]) {
    const int in_size = 15, center = 7 * in_size + 7;
    for (int i = 0; i < 8; i++) {
        for (int j = 0; j < 8; j++) {
            int i0 = center - i * in_size - j, i1 = center - i * in_size + j;
            int i2 = center + i * in_size - j, i3 = center + i * in_size + j;
            sym[0][i * 8 + j] = 0.25 * (k[i0] + k[i1] + k[i2] + k[i3]);
            if (j > 0) sym[1][i * 8 + j - 1] = 0.25 * (-k[i0] + k[i1] - k[i2] + k[i3]);
            if (i > 0) sym[2][(i - 1) * 8 + j] = 0.25 * (-k[i0] - k[i1] + k[i2] + k[i3]);
            if (i > 0 && j > 0) sym[3][(i - 1) * 8 + (j - 1)] = 0.25 * (-k[i0] + k[i1] - k[i2] + k[i3]);
        }
        sym[1][i * 8 + 7] = 0.0;
        sym[2][7 * 8 + i] = 0.0;
        sym[3][i * 8 + 7] = 0.0;
        sym[3][7 * 8 + i] = 0.0;
    }
}

/* ID:4792-4851 clt_fill_coord_corr over one colour's kernel grid: least-squares slopes of the
 * centres over the 3x3 neighbourhood; ext: [ky][kx][8] */
static void fill_coord_corr(int kern_step, int kx, int ky, double* ext) {
    double* dst = malloc(sizeof(double) * kx * ky * 4);
    for (int ty = 0; ty < ky; ty++)
        for (int tx = 0; tx < kx; tx++) {
            double s0 = 0, sx = 0, sx2 = 0, sy = 0, sy2 = 0, sz = 0, sxz = 0, syz = 0, sw = 0, sxw = 0, syw = 0;
            for (int dty = -1; dty < 2; dty++) {
                int y = ty + dty;
                if (y < 0 || y >= ky) continue;
                for (int dtx = -1; dtx < 2; dtx++) {
                    int x = tx + dtx;
                    if (x < 0 || x >= kx) continue;
                    const double* e = ext + (y * kx + x) * 8;
                    s0 += 1;
                    sx += dtx;
                    sx2 += dtx * dtx;
                    sy += dty;
                    sy2 += dty * dty;
                    sz += e[2];
                    sxz += dtx * e[2];
                    syz += dty * e[2];
                    sw += e[3];
                    sxw += dtx * e[3];
                    syw += dty * e[3];
                }
            }
            double denom_x = (sx2 * s0 - sx * sx) * kern_step;
            double denom_y = (sy2 * s0 - sy * sy) * kern_step;
            double* d = dst + (ty * kx + tx) * 4;
            d[0] = (sxz * s0 - sz * sx) / denom_x;
            d[1] = (syz * s0 - sz * sy) / denom_y;
            d[2] = (sxw * s0 - sw * sx) / denom_x;
            d[3] = (syw * s0 - sw * sy) / denom_y;
        }
    for (int t = 0; t < kx * ky; t++)
        for (int i = 0; i < 4; i++) ext[t * 8 + 4 + i] = dst[t * 4 + i];
    free(dst);
}

/* calculateCLTKernel (QC:433-668) for one camera.  raw: the kernel stack [colors][ky*ks][kx*ks]
 * (extractOneKernel QC:1071-1089 takes kernel (tx, ty) at ((ty*kx*ks + tx)*ks)); sensor_xy:
 * [ky][kx][2] sensor correction at the kernel centres (QC:603-610, offsetKernelSensor ID:4781-4791
 * by its negative) or NULL.  Outputs in the upload layout of setConvolutionKernels (GTP:381-435):
 * kern [ky][kx][colors][4][v][h], offs [ky][kx][colors][8], rounded to float. */
EXPORT int orc_prepare_kernels(const float* raw, int ks, int kx, int ky, int colors, int kern_step,
                               const float* sensor_xy, int norm_kern, float* kern_out, float* offs_out) {
    tables();
    if (ks < 32 || (ks & 1) || kx < 1 || ky < 1 || colors < 1) return CLT_E_ARG;
    double window[64];
    for (int i = 0; i < 8; i++)   /* QC:466-474 norm_sym_weights */
        for (int j = 0; j < 8; j++) window[i * 8 + j] = cos(M_PI * i / 16) * cos(M_PI * j / 16);
    double* k = malloc(sizeof(double) * ks * ks);
    double* ext = malloc(sizeof(double) * ky * kx * 8);
    const size_t W = (size_t)kx * ks;
    for (int col = 0; col < colors; col++) {
        const float* plane = raw + (size_t)col * ky * ks * W;
        for (int ty = 0; ty < ky; ty++)
            for (int tx = 0; tx < kx; tx++) {
                const size_t base = ((size_t)ty * W + tx) * ks;
                for (int i = 0; i < ks; i++)
                    for (int j = 0; j < ks; j++) k[i * ks + j] = plane[base + i * W + j];
                double kc[225 + 4], sym[4][64];
                convert_double_kernel(k, ks, kc);
                if (norm_kern) normalize_kernel(kc, window);
                symmetrize_kernel(kc, sym);
                float* ko = kern_out + (((size_t)ty * kx + tx) * colors + col) * 256;
                for (int q = 0; q < 4; q++) {   /* ID:4725-4734 clt_dtt3_kernel, output [h][v] */
                    double y[64];
                    real xr[64], yr[64];
                    for (int i = 0; i < 64; i++) xr[i] = (real)sym[q][i];
                    dttt_iiie(xr, q, yr);
                    for (int i = 0; i < 64; i++) y[i] = yr[i];
                    for (int v = 0; v < 8; v++)
                        for (int hh = 0; hh < 8; hh++) ko[q * 64 + v * 8 + hh] = (float)y[hh * 8 + v];
                }
                double* e = ext + (ty * kx + tx) * 8;
                for (int i = 0; i < 4; i++) e[i] = kc[225 + i];
                for (int i = 4; i < 8; i++) e[i] = 0.0;
                if (sensor_xy) {
                    const double dx = -(double)sensor_xy[(ty * kx + tx) * 2], dy = -(double)sensor_xy[(ty * kx + tx) * 2 + 1];
                    e[2] += dx;
                    e[3] += dy;
                    e[0] += dx;
                    e[1] += dy;
                }
            }
        fill_coord_corr(kern_step, kx, ky, ext);
        for (int t = 0; t < kx * ky; t++)
            for (int i = 0; i < 8; i++) offs_out[((size_t)t * colors + col) * 8 + i] = (float)ext[t * 8 + i];
    }
    free(k);
    free(ext);
    return 0;
}
