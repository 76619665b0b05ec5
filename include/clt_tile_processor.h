/*
 * clt_tile_processor.h -- C ABI of the MI355X (gfx950) Complex Lapped Transform tile processor.
 *
 * This is the drop-in boundary for the GPU path of imagej-elphel's tile processor.  Every entry
 * point below replaces one method of the reference's JCuda wrapper
 *   src/main/java/com/elphel/imagej/gpu/GPUTileProcessor.java            (GTP below)
 * or adds the GPU form of a CPU-only stage of
 *   src/main/java/com/elphel/imagej/tileprocessor/ImageDtt.java          (ID below)
 *   src/main/java/com/elphel/imagej/tileprocessor/Correlation2d.java     (C2 below)
 * The semantics of every result are those of the reference CPU path (ImageDtt); the memory layouts
 * of the host buffers are those the reference GPU path exchanges with Java (GTP), unless noted.
 *
 * Conventions
 *   - All functions return 0 on success, a negative CLT_E_* code on failure; the message of the
 *     last failure of a handle is returned by clt_last_error() (GTP throws CudaException instead).
 *   - Plain C types only; host buffers are borrowed for the duration of a call.
 *   - A handle is not thread safe.  Every clt_exec_* call returns after the device work finished
 *     (GTP:550-557 synchronises the context after every launch) unless clt_set_async(h, 1).
 *   - Colours of a Bayer frame are numbered as in the reference: 0 = red, 1 = blue, 2 = green
 *     (EyesisCorrections.java:2192, 2219).  A monochrome frame has one colour (0).
 *   - 8x8 blocks of CLT coefficients exchanged with the host are row-major [vertical][horizontal]
 *     (the layout of the reference GPU, GTP:406-408, TileProcessor.cuh:1065-1095).
 */
#ifndef CLT_TILE_PROCESSOR_H
#define CLT_TILE_PROCESSOR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLT_DTT_SIZE        8    /* transform size (ECP:2519 transform_size)          */
#define CLT_TILE_COEFFS   256    /* 4 quadrants x 8 x 8 coefficients per tile/colour   */
#define CLT_CORR_SIZE      15    /* 2*DTT_SIZE-1 (C2:134)                             */
#define CLT_CORR_LEN      225
#define CLT_MAX_CAMS       16
#define CLT_MAX_PAIRS     120
#define CLT_DISP_LAYERS    12    /* per-tile disparity record, see CLT_DISP_*          */
#define CLT_TEX_STATS      16    /* per-tile texture statistics of a quad (4*num_cams)  */

enum {
    CLT_OK            =  0,
    CLT_E_ARG         = -1,  /* invalid argument (null pointer, size out of range)            */
    CLT_E_STATE       = -2,  /* required input not set yet (kernels, images, tasks)            */
    CLT_E_HIP         = -3,  /* HIP runtime error                                              */
    CLT_E_NODEVICE    = -4,  /* no GPU visible                                                 */
    CLT_E_NOMEM       = -5,  /* device allocation failed                                       */
    CLT_E_UNSUPPORTED = -6   /* configuration not supported by this entry point                */
};

/* Per-tile disparity record written by clt_exec_corr2d (ID:80-86 DISPARITY_TITLES subset). */
enum {
    CLT_DISP_INT        = 0, /* -argmax on the combined strip axis          (ID:2197)           */
    CLT_DISP_CM         = 1, /* centre-of-mass disparity                    (ID:2244, C2:936)   */
    CLT_DISP_STRENGTH   = 2, /* combined strip value at the integer maximum (ID:2196)           */
    CLT_DISP_HOR        = 3, /* horizontal pairs, parabola                  (ID:2216-2225)      */
    CLT_DISP_HOR_STR    = 4,
    CLT_DISP_VERT       = 5, /* vertical pairs, parabola                    (ID:2227-2237)      */
    CLT_DISP_VERT_STR   = 6,
    CLT_DISP_CM_WEIGHT  = 7, /* total weight of the centre-of-mass window   (C2:1025)           */
    /* With the per-tile correlation LMA enabled (clt_corr_lma_params.enable, the reference
     * default), CLT_DISP_CM and CLT_DISP_STRENGTH hold the final values of ID:2254-2336: the LMA
     * disparity/strength when mix_corr_poly, then the foreground correction.  Without it they
     * hold the centre of mass and the integer-maximum strength. */
    CLT_DISP_POLY       = 8, /* LMA disparity (ID:2266 DISPARITY_INDEX_POLY), NaN if not fitted */
    CLT_DISP_POLY_STR   = 9, /* LMA strength (Correlations2dLMA.getDisparityStrength)           */
    CLT_DISP_CM_RAW     = 10,/* centre-of-mass disparity before the LMA                         */
    CLT_DISP_STR_RAW    = 11 /* integer-maximum strength before the LMA                         */
};

/* Frame geometry: fixed for the life of a handle (the reference injects the same values as
 * #defines before NVRTC compilation, GTP:80-97, GTP:262-276). */
typedef struct clt_geom {
    int32_t width;          /* pixels, multiple of 8 (IMG_WIDTH)                              */
    int32_t height;         /* pixels, multiple of 8 (IMG_HEIGHT)                             */
    int32_t num_cams;       /* 1..16 (NUM_CAMS)                                               */
    int32_t num_colors;     /* 3 = Bayer mosaic (R,B,G), 1 = monochrome (NUM_COLORS)          */
    int32_t red_row;        /* Bayer phase of red: row parity (reference: 0, TPC:113)         */
    int32_t red_col;        /* Bayer phase of red: column parity (reference: 1, TPC:114)      */
    int32_t kernels_hor;    /* aberration kernel grid width  (KERNELS_HOR  = W/16+2)          */
    int32_t kernels_vert;   /* aberration kernel grid height (KERNELS_VERT = H/16+2)          */
    int32_t kernel_step;    /* pixels between kernel centres, 16 (ECP:2545)                  */
    int32_t max_tasks;      /* capacity of the task list; 0 = tilesX * tilesY                 */
} clt_geom;

/* One correlation pair: cameras and orientation (C2:52-71 PAIRS).  dir: 0 horizontal,
 * 1 vertical, 2 main diagonal, 3 other diagonal, -1 any other direction (only the 15x15
 * correlation is produced for such pairs).  scale: baseline scale (1 for a quad camera). */
typedef struct clt_pair {
    int32_t cam1, cam2, dir, scale;
} clt_pair;

/* Correlation parameters (ImageDttParameters IDP:31-112, CLTParameters ECP:2529-2573). */
typedef struct clt_corr_params {
    float   fat_zero;        /* phase-correlation fat zero, 0.0            (ECP:2529)          */
    float   corr_red;        /* red-to-green weight, 0.5                   (ECP:2557)          */
    float   corr_blue;       /* blue-to-green weight, 0.2                  (ECP:2558)          */
    float   corr_offset;     /* strip product offset, 0.1                  (IDP:67)            */
    int32_t twice_diagonal;  /* 1                                          (IDP:68)            */
    float   min_corr;        /* minimal integer maximum, 0.02              (IDP:69)            */
    int32_t strip_height;    /* rows of the combined strip, 9              (IDP:73)            */
    float   enhortho_width;  /* ortho notch width, 2                       (IDP:44)            */
    float   enhortho_scale;  /* ortho notch scale, 0.0                     (IDP:45)            */
} clt_corr_params;

/* Texture parameters (ECP:2584-2594, ID:1491-1555). */
typedef struct clt_tex_params {
    float   min_shot;          /* 10.0   */
    float   scale_shot;        /* 3.0    */
    float   diff_sigma;        /* 5.0    */
    int32_t dust_remove;       /* 1      */
    float   corr_magic_scale;  /* 0.85   (ECP corr_magic_scale)                               */
    int32_t keep_weights;      /* 1: 12 planes per tile (R,B,G,A,w0..w3,rms_r,rms_b,rms_g,rms),
                                  0: 4 planes (R,B,G,A)                                        */
    float   port_offsets[CLT_MAX_CAMS][2]; /* per camera {x,y} offset (ID:1648-1652); default
                                            * for other than 4 cameras: a circle of radius 0.5 */
} clt_tex_params;

/* Per-tile correlation LMA and foreground correction (ImageDttParameters, IDP:32-110): the fit of
 * Correlations2dLMA (C2:1754-1900, C2:2029-2338) run on the combined pairs when the strength
 * exceeds min_poly_strength (ID:2254-2288), then corr4dirsLMA + foregroundCorrect
 * (C2:1591-1750, ID:2288-2336).  Doubles, as in the reference. */
typedef struct clt_corr_lma_params {
    int32_t enable;              /* 1: clt_exec_corr2d runs the chain (the reference always does)  */
    int32_t mix_corr_poly;       /* 1: LMA result replaces CM and strength        (IDP:32)         */
    double  min_poly_strength;   /* 0.2                                           (IDP:33)         */
    double  ortho_vasw_pwr;      /* 2.0: sample weight *= |value|^pwr             (IDP:46)         */
    int32_t cnvx_hwnd_size;      /* 4 (at most 4): half window of the convex cells (IDP:76)        */
    int32_t cnvx_add3x3;         /* 1                                             (IDP:78)         */
    double  cnvx_weight;         /* 0.5: weight of non-convex selected cells      (IDP:77)         */
    int32_t lma_adjust_wm;       /* 1                                             (IDP:95)         */
    int32_t lma_adjust_wy;       /* 0                                             (IDP:96)         */
    int32_t lma_adjust_wxy;      /* 1                                             (IDP:97)         */
    int32_t lma_adjust_ag;       /* 1                                             (IDP:98)         */
    double  lma_half_width;      /* 2.0                                           (IDP:100)        */
    double  lma_cost_wy;         /* 0.1                                           (IDP:101)        */
    double  lma_cost_wxy;        /* 0.1                                           (IDP:102)        */
    double  lma_lambda_initial;  /* 0.1                                           (IDP:104)        */
    double  lma_lambda_scale_good; /* 0.5                                         (IDP:105)        */
    double  lma_lambda_scale_bad;  /* 8.0                                         (IDP:106)        */
    double  lma_lambda_max;      /* 100.0                                         (IDP:107)        */
    double  lma_rms_diff;        /* 0.001                                         (IDP:108)        */
    int32_t lma_num_iter;        /* 20                                            (IDP:109)        */
    int32_t fo_correct;          /* 1                                             (IDP:54)         */
    int32_t fo_far;              /* 0                                             (IDP:55)         */
    int32_t fo_ortho;            /* 1                                             (IDP:61)         */
    double  fo_min_strength;     /* 0.1                                           (IDP:56)         */
    double  fo_min_eff;          /* 0.25                                          (IDP:57)         */
    double  fo_min_eff_ratio;    /* 0.65                                          (IDP:58)         */
    double  fo_max_hwidth;       /* 3.0                                           (IDP:59)         */
    double  fo_min_diff;         /* 0.2                                           (IDP:60)         */
    double  fo_overcorrection;   /* 0.0                                           (IDP:62)         */
    double  fo_lim_overcorr;     /* 2.0                                           (IDP:63)         */
} clt_corr_lma_params;

/* Two-quad rig (ImageDtt.clt_bi_quad, ID:8731-9504): BiQuadParameters (BQP) and the
 * ImageDttParameters / CLTParameters fields the inter-camera stage reads beyond the ones above. */
typedef struct clt_rig_params {
    int32_t use_poly;             /* LMA replaces the centre-of-mass disparity, 1    (BQP:33)     */
    int32_t use_xy_poly;          /* dx, dy by the 2-D polynomial (else CM), 1       (BQP:34)     */
    int32_t no_int_x0;            /* inter CM around 0, not the integer max, 1       (BQP:32)     */
    int32_t notch_mode;           /* pole detection of the inter correlation, 0      (TQ:312)     */
    int32_t lt_rad;               /* low-texture averaging radius in tiles, 0        (BQP:83)     */
    int32_t corr_strip_notch;     /* rows of the notch window, 11                    (IDP:81)     */
    double  corr_notch_hwidth;    /* 2.0                                             (IDP:82)     */
    double  corr_notch_blur;      /* 2.0                                             (IDP:83)     */
    double  cm_max_normalization; /* 0.55                                            (IDP:112)    */
    double  max_corr_radius;      /* single2dCM radius, 3.9                          (ECP:2566)   */
} clt_rig_params;

/* Per-task rig record written by clt_exec_corr2d_rig: the BIDISPARITY layers of clt_bi_quad
 * (ID:89-114); the main quad is cameras 0..3, the auxiliary quad cameras 4..7. */
enum {
    CLT_BI_DISP_FULL = 0, CLT_BI_DISP_HOR, CLT_BI_DISP_VERT, CLT_BI_DISP_DIAGM, CLT_BI_DISP_DIAGO,
    CLT_BI_ADISP_FULL = 5, CLT_BI_ADISP_HOR, CLT_BI_ADISP_VERT, CLT_BI_ADISP_DIAGM, CLT_BI_ADISP_DIAGO,
    CLT_BI_DISP_CROSS = 10, CLT_BI_DISP_CROSS_DX, CLT_BI_DISP_CROSS_DY,
    CLT_BI_STR_FULL = 13, CLT_BI_STR_HOR, CLT_BI_STR_VERT, CLT_BI_STR_DIAGM, CLT_BI_STR_DIAGO,
    CLT_BI_ASTR_FULL = 18, CLT_BI_ASTR_HOR, CLT_BI_ASTR_VERT, CLT_BI_ASTR_DIAGM, CLT_BI_ASTR_DIAGO,
    CLT_BI_STR_CROSS = 23, CLT_BI_STR_ALL = 24, CLT_BI_TARGET = 25,
    CLT_BI_LAYERS = 26
};

typedef struct clt_handle clt_handle;

/* ---- life cycle: GPUTileProcessor() (GTP:243-360); the reference never frees ---------------- */
int  clt_create(const clt_geom* geom, clt_handle** out);
int  clt_destroy(clt_handle* h);
const char* clt_last_error(const clt_handle* h);
int  clt_device_count(void);                              /* 0 when no GPU is visible          */

/* ---- inputs ----------------------------------------------------------------------------------- */
/* setConvolutionKernel(float[] kernel, float[] kernel_offsets, int ncam) (GTP:373-379).
 * kernels: [kernels_vert][kernels_hor][num_colors][4][8][8] floats, each 8x8 [v][h];
 * offsets: [kernels_vert][kernels_hor][num_colors][8] CltExtra records (TPC:139-148). */
int  clt_set_kernels(clt_handle* h, int32_t cam, const float* kernels, const float* offsets);
/* Drop the kernels: convert without deconvolution (ID:4896 use_kernels == false). */
int  clt_clear_kernels(clt_handle* h);
/* calculateCLTKernel(sensor, kernelStack, kernelSize, clt_parameters, ...) (QC:433-668) for one
 * camera, computed on the device and installed as by clt_set_kernels.  raw: the kernel stack
 * [num_colors][kernels_vert * ks][kernels_hor * ks] floats (one slice per colour, kernel (tx, ty)
 * at rows ty*ks.., columns tx*ks.., extractOneKernel QC:1071-1089); ks: raw kernel size (64,
 * even, 32..256); sensor_xy: [kernels_vert][kernels_hor][2] floats, the sensor correction vector
 * at each kernel centre (interpolateCorrectionVector at ((tx - 0.5) kernel_step, (ty - 0.5)
 * kernel_step), QC:603-610) or NULL for none; norm_kern: clt_parameters.norm_kern (QC:464). */
int  clt_prepare_kernels(clt_handle* h, int32_t cam, const float* raw, int32_t ks,
                         const float* sensor_xy, int32_t norm_kern);
/* The installed kernels of one camera in the layout clt_set_kernels takes (GTP:381-435), e.g. to
 * save them (saveFloatKernels TQ:1076-1177). */
int  clt_get_kernels(clt_handle* h, int32_t cam, float* kernels, float* offsets);
/* flattenCLTKernels (QC:670-745) / extractCLTKernels (QC:790-857) between the clt_set_kernels
 * layout and the flat per-colour image [num_colors][17 ky][16 kx] (each kernel tile: CC | SC over
 * CS | SS as 8x8 blocks in the reference's CPU orientation, then the 8 CltExtra values in the
 * 17th row).  Host only; no handle. */
int  clt_kernels_to_flat(const float* kernels, const float* offsets, int32_t kx, int32_t ky,
                         int32_t colors, float* flat);
int  clt_kernels_from_flat(const float* flat, int32_t kx, int32_t ky, int32_t colors,
                           float* kernels, float* offsets);
/* setBayerImage(float[] bayer_image, int ncam) (GTP:437-450).  image: [height][stride] floats,
 * colours of a Bayer frame interleaved as a mosaic (green already scaled by the caller,
 * TQ:2062-2067), stride >= width. */
int  clt_set_image(clt_handle* h, int32_t cam, const float* image, int32_t stride);
/* setTasks(TpTask[], boolean) (GTP:363-371).  tasks: num_tasks records of (2 + 2*num_cams)
 * 32-bit words: int task (0 = skip; bits 0..7 = ports used by textures), int txy (tx | ty<<16),
 * float xy[num_cams][2] (port tile centres, GTP:132-166). */
int  clt_set_tasks(clt_handle* h, const void* tasks, int32_t num_tasks);
/* setLpfRbg(float,float,float) (GTP:780-812): LPF applied by convert for colours in lpf_mask.
 * setCltLpf(double) (GTP:814-848) is clt_lpf_coeffs below. */
int  clt_set_lpf_rbg(clt_handle* h, float sigma_r, float sigma_b, float sigma_g);
/* Correlation/texture/imclt low-pass filter sigma (corr_sigma, ID:1662-1690, ID:4178-4271). */
int  clt_set_corr_sigma(clt_handle* h, float sigma);
int  clt_set_pairs(clt_handle* h, const clt_pair* pairs, int32_t num_pairs);
int  clt_set_corr_params(clt_handle* h, const clt_corr_params* p);
int  clt_set_tex_params(clt_handle* h, const clt_tex_params* p);
int  clt_set_corr_lma_params(clt_handle* h, const clt_corr_lma_params* p);
/* 64 LPF multipliers for sigma, [h][v] order (GTP:814-848 setCltLpf, then dttt_iiie * 2n). */
int  clt_lpf_coeffs(double sigma, float out[64]);
/* Defaults of the reference parameter classes (IDP, ECP); quad camera pairs C2:52-58. */
void clt_default_corr_params(clt_corr_params* p);
void clt_default_tex_params(clt_tex_params* p, int32_t num_cams);
void clt_default_corr_lma_params(clt_corr_lma_params* p);
void clt_default_rig_params(clt_rig_params* p);
/* BiQuadParameters of the two-quad rig path (clt_exec_corr2d_rig). */
int  clt_set_rig_params(clt_handle* h, const clt_rig_params* p);
int  clt_default_quad_pairs(clt_pair out[6]);

/* ---- kernels ---------------------------------------------------------------------------------- */
/* execConverCorrectTiles() (GTP:530-558): convert_direct over the task list. */
int  clt_exec_convert(clt_handle* h, int32_t lpf_mask);
/* execImcltRbg() (GTP:560-597): inverse CLT with overlap-add of every converted tile of the last
 * convert; apply_lpf != 0 multiplies by the corr_sigma LPF first (clt_lpf, ID:4178-4271). */
int  clt_exec_imclt(clt_handle* h, int32_t apply_lpf);
/* Phase correlation of every pair for every task (C2:197-389), combined strip (C2:556-790),
 * per-tile disparity (ID:2186-2238) and, when enabled, the per-tile LMA chain (ID:2254-2336) over
 * the pairs with a known orientation (at most 2 compatible groups, C2:72-78). */
int  clt_exec_corr2d(clt_handle* h);
/* clt_exec_convert(h, lpf_mask) then clt_exec_corr2d(h) with the same outputs (GTP:530-558 then
 * execCorrelate2D, the order of the reference's frame loop): in the quad layout one kernel, the CLT
 * of each task handed from the conversion to the correlation on chip (it is still written for the
 * textures and imclt); otherwise the two calls. */
int  clt_exec_convert_corr2d(clt_handle* h, int32_t lpf_mask);
/* The correlation part of clt_bi_quad(...) (ID:8731-9504) for a two-quad rig held as one 8-camera
 * frame: cameras 0..3 the main quad, 4..7 the auxiliary quad (tasks carry TpTask.xy then xy_aux,
 * GTP:132-166).  Per task: tileCorrs (ID:7581-7740) of each quad over the quad pair table
 * (correlate2D, the LMA over all pairs replacing the centre of mass when use_poly, the four
 * one-direction LMA fits), correlateInterCamerasFD (C2:401-448) of the two quads' mean CLT,
 * tileInterCamCorrs (ID:7370-7560), with lt_rad > 0 the low-texture pass (ID:9404-9480), and the
 * BIDISPARITY record (CLT_BI_*).  Uses the correlation, LMA and rig parameters and corr_sigma;
 * target_disparity: num_tasks floats for CLT_BI_TARGET or NULL (NaN). */
int  clt_exec_corr2d_rig(clt_handle* h, const float* target_disparity);
/* generateTextureTiles of both quads after clt_exec_corr2d_rig (ID:9360-9398): the 4 ports of each
 * quad with the quad port offsets (ID:8832-8836) and the quad's tileCorrs disparity
 * (CLT_BI_DISP_FULL / CLT_BI_ADISP_FULL) as the extra disparity; the texture parameters as set
 * (the port offsets excepted). */
int  clt_exec_textures_rig(clt_handle* h);
/* Per-tile textures (ID:2662-2812).  extra_disparity: num_tasks floats or NULL (ID:2663-2677). */
int  clt_exec_textures(clt_handle* h, const float* extra_disparity);
/* No single reference counterpart: clt_exec_textures(h, extra_disparity) followed by
 * clt_exec_imclt(h, apply_lpf) (ID:2662-2812 and execImcltRbg GTP:560-597), with the same outputs.
 * When extra_disparity is NULL, apply_lpf != 0, there are at most 5 cameras and every task has
 * non-zero flags and a tile of its own, one pass computes each tile's inverse CLT once for both
 * outputs; otherwise the two calls run. */
int  clt_exec_textures_imclt(clt_handle* h, const float* extra_disparity, int32_t apply_lpf);
/* Overlap-add of the texture tiles into full-frame planes (combineRGBATiles, ID:3858-3989,
 * overlap = true, sharp_alpha = false). */
int  clt_exec_combine_textures(clt_handle* h);
/* combineRGBATiles(texture_tiles, 8, overlap, sharp_alpha, ...) (ID:3855-3991) of the last textures:
 * overlap != 0: [planes][height][width], the 16x16 tiles overlap-added at stride 8, with sharp_alpha
 * the alpha plane from the centre 8x8 of each tile only; overlap == 0: the [planes][2 height]
 * [2 width] mosaic of whole tiles.  A tile position holds the last task on it with a non-zero task
 * word (texture_tiles[tileY][tileX]); positions without one stay 0. */
int  clt_exec_combine_rgba_tiles(clt_handle* h, int32_t overlap, int32_t sharp_alpha);

/* ---- outputs ---------------------------------------------------------------------------------- */
/* getRBG(int ncam) (GTP:599-626): [num_colors][height+8][width+8] floats (uncropped; the
 * reference CPU image is this one cropped by 4 pixels, ID:3617). */
int  clt_get_rbg(clt_handle* h, int32_t cam, float* out);
/* CLT of one camera: [tilesY*tilesX][num_colors][4][8][8] floats, 8x8 as [v][h]. */
int  clt_get_clt(clt_handle* h, int32_t cam, float* out);
/* corr: [num_tasks][num_pairs][225]; combo: [num_tasks][strip_height*15] (may be NULL);
 * disparity: [num_tasks][CLT_DISP_LAYERS] (may be NULL). */
int  clt_get_corr2d(clt_handle* h, float* corr, float* combo, float* disparity);
/* Correlation-peak offsets of the 6 quad pairs after clt_exec_corr2d: the per-pair measurement of
 * the reference's extrinsic path (clt_mismatch, ID:2342-2420, mismatchPairsCM C2:1336-1430), which
 * the coarse stage of clt_lma_extrinsics solves for.  offsets [num_tasks][6][3]: dx, dy of the
 * centre of mass of the positive correlation values within radius (1..7) of the pair's integer
 * maximum (image axes, pixels from the centre: the excess of the pair's relative port offset
 * xy_c1 - xy_c2 over the aligned one; NaN where no value there is positive) and the correlation
 * at the maximum.  Tasks without work: NaN, NaN, 0.  The reference centres its window at the tile's
 * residual disparity (xcenter = -disparity) and reports offsets beyond it; here the pairs are
 * aligned at the task disparity and the disparity is solved for with the extrinsics. */
int  clt_get_corr_offsets(clt_handle* h, double radius, float* offsets);
/* Results of clt_exec_corr2d_rig, each may be NULL: bimap [num_tasks][CLT_BI_LAYERS] (the
 * disparity_bimap layers of clt_bi_quad); inter [num_tasks][225] inter-camera correlations
 * (clt_corr_combo[0], before the low-texture averaging); corr_main / corr_aux [num_tasks][6][225]
 * pair correlations; disp_main / disp_aux [num_tasks][CLT_DISP_LAYERS] correlate2D records of
 * each quad after its LMA (CM = the tileCorrs disparity, STRENGTH its strength). */
int  clt_get_corr2d_rig(clt_handle* h, float* bimap, float* inter, float* corr_main, float* corr_aux,
                        float* disp_main, float* disp_aux);
/* Results of clt_exec_textures_rig, each may be NULL: tiles [num_tasks][planes][256] with planes =
 * 12 when keep_weights (RGBA, 4 port weights, rms r,b,g,combined), else 4; stats [num_tasks][16]
 * as clt_get_textures for 4 cameras. */
int  clt_get_textures_rig(clt_handle* h, float* tiles_main, float* tiles_aux, float* stats_main,
                          float* stats_aux);
/* tiles: [num_tasks][planes][256] with planes = 4 + num_cams + 4 when keep_weights (RGBA, one
 * weight per port, rms r,b,g,combined), else 4 (RGBA); stats:
 * [num_tasks][4*num_cams] = max_diff[num_cams] (ID:2802-2806) then the mean tones
 * r[0..C-1], b[0..C-1], g[0..C-1] (ID:2807-2811), or NULL. */
int  clt_get_textures(clt_handle* h, float* tiles, float* stats);
/* The planes of the last clt_exec_combine_textures / clt_exec_combine_rgba_tiles:
 * [planes][height][width], or [planes][2 height][2 width] for the mosaic. */
int  clt_get_combined_textures(clt_handle* h, float* out);

/* ---- port geometry on the device ---------------------------------------------------------------
 * The reference computes the port tile centres of every task on the host in double precision
 * (getTileSubcamOffsets, GTP:686-778, calling GeometryCorrection.getPortsCoordinatesAndDerivatives,
 * GC:2159-2314) before setTasks.  Here the same model runs on the GPU in double precision: undo the
 * common radial distortion of the tile centre, subtract disparity * rXY, rotate/zoom per sensor
 * (CorrVector, GC:1043-1147), project, re-apply the radial distortion, add the sensor centre. */
#define CLT_EXTRINSICS 13   /* CorrVector: tilt0..2, azimuth0..2, roll0..3, zoom0..2 (GC:1004-1011) */
typedef struct clt_port_geometry {
    int32_t num_sensors;        /* 4: the CorrVector model is that of a quad camera               */
    int32_t width, height;      /* pixelCorrectionWidth, pixelCorrectionHeight (GC:45-46)         */
    int32_t reserved;
    double  focal_length;       /* mm (GC:47)                                                     */
    double  pixel_size;         /* um (GC:48)                                                     */
    double  distortion_radius;  /* mm (GC:49)                                                     */
    double  distortion[7];      /* distortionC, B, A, A5, A6, A7, A8: the r^2..r^8 terms (GC:2188) */
    double  rxy[4][2];          /* port offsets in disparity units (GC:77 rXY)                    */
    double  pxy0[4][2];         /* sensor centres, pixels (GC:72 pXY0)                            */
    double  roll_deg[4];        /* factory rolls, degrees (GC:71 roll)                            */
    double  extrinsic[CLT_EXTRINSICS];   /* CorrVector.vector                                     */
} clt_port_geometry;

/* GeometryCorrection.RigOffset (GC:256-290) of the auxiliary quad of a two-quad rig, and the
 * disparityRadius of each camera (GC:81, GC:1757: sqrt(2) * cameraRadius, mm). */
typedef struct clt_rig_offset {
    double  baseline;           /* mm between the camera centres, 1256 (GC:268)                   */
    double  aux_angle;          /* radians, 0: the auxiliary camera to the right (GC:269)         */
    double  aux_z;              /* mm towards the scene; not used by the port centres (GC:271)   */
    double  aux_azimuth;        /* radians, positive looks right (GC:272)                         */
    double  aux_tilt;           /* radians, positive looks up (GC:273)                            */
    double  aux_roll;           /* radians, positive clockwise (GC:274)                           */
    double  aux_zoom;           /* relative zoom - 1 (GC:275)                                     */
    double  main_disparity_radius, aux_disparity_radius;   /* mm                                 */
} clt_rig_offset;

/* Zero distortion, zero extrinsics, quad rXY {(-1/2,-1/2),(1/2,-1/2),(-1/2,1/2),(1/2,1/2)}
 * (GC:78), sensor centres at the frame centre, reference focal length/pixel/radius (GC:37-39). */
void clt_default_port_geometry(clt_port_geometry* p, int32_t width, int32_t height);
/* Rotation matrices, their derivatives and the reverse distortion table (calcReverseDistortionTable,
 * GC:2707-2756) are prepared on the host and uploaded; needs a handle of 4 cameras. */
int  clt_set_port_geometry(clt_handle* h, const clt_port_geometry* p);
/* The port geometry of a two-quad rig on an 8-camera handle, for clt_set_tasks_geometry: the main
 * quad (sensors 0..3) as clt_set_port_geometry; the auxiliary quad (sensors 4..7) as clt_bi_quad
 * computes it (ID:8970-8993): getPortsCoordinatesAndDerivatives(gc_main, use_rig_offsets = true,
 * aux CorrVector rotations times RigOffset.getRotMatrix, ...) at the disparity scaled by
 * aux_disparity_radius / main_disparity_radius.  No derivatives (clt_get_port_derivatives). */
int  clt_set_port_geometry_rig(clt_handle* h, const clt_port_geometry* main, const clt_port_geometry* aux,
                               const clt_rig_offset* rig);
/* getTileSubcamOffsets + setTasks: task words, txy = tx | ty << 16 and target disparities of
 * num_tasks tiles; the tile centre is (8 tx + 4, 8 ty + 4) (GTP:724-725).  Records of tasks with a
 * zero task word keep zero centres. */
int  clt_set_tasks_geometry(clt_handle* h, const int32_t* task, const int32_t* txy, const float* disparity,
                            int32_t num_tasks);
/* Copy of the current task records, [num_tasks][2 + 2*num_cams] 32-bit words. */
int  clt_get_tasks(clt_handle* h, void* tasks);
/* Derivatives of the port centres of the current tasks: [num_tasks][2*num_sensors][14] doubles,
 * rows x0, y0, x1, y1, ...; columns 0..12 by the extrinsic parameters (pXYderiv, GC:2245-2306),
 * column 13 by the disparity.  Computed for tasks set by clt_set_tasks_geometry. */
int  clt_get_port_derivatives(clt_handle* h, double* out);

/* ---- correlation-peak loss and its adjoint (no reference counterpart) ---------------------------
 * Loss of a task: l = -sum_p C_p, C_p the zero-offset (centre) value of pair p's 15x15 phase
 * correlation (C2:275-389 with the unfold of DR:443-489).  Gradient by the port centres of the task
 * through convert_direct (ID:4867-5584) by a hand-written forward/reverse pass, no CLT in HBM.
 * Quad Bayer handle, the 6 reference pairs, fat_zero 0.  Per task CLT_CORR_ADJ floats:
 * [0] loss, [1..6] C_p, [7..12] and [13..18] the curvatures -d2 C_p / d tau^2 under a relative
 * shift of the pair along x and y, [19..26] dl/d(x0, y0, x1, y1, x2, y2, x3, y3), rest 0.
 * A task with task word 0, or with a window of any camera outside the frame, gets NaN loss and
 * zero gradient (and is left out of the sums below). */
#define CLT_CORR_ADJ 32
int  clt_exec_corr_adjoint(clt_handle* h);
int  clt_get_corr_adjoint(clt_handle* h, float* out);

/* Gradient of the summed loss of the current tasks (set by clt_set_tasks_geometry) by the 13
 * extrinsic parameters (CorrVector, GC:1004-1011) and by each task's disparity, through the port
 * geometry derivatives (GC:2235-2306).  Any output may be NULL. */
int  clt_exec_geometry_adjoint(clt_handle* h, double* grad_extrinsic, float* grad_disparity, double* loss);

/* No counterpart in the reference.  clt_exec_geometry_adjoint without any device allocation, for
 * callers that cannot allocate (code under HIP graph capture, tight serving or LMA loops, harnesses
 * that forbid hipMalloc): the same kernels and outputs, with the per-task adjoint records, per-task
 * state, per-block partial systems and their sums in one static device array of
 * CLT_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends).  At most CLT_NOALLOC_MAX_TASKS
 * tasks (a full 2592x1936 frame is 78,408); more fail with CLT_E_ARG.  The handle's inputs come from
 * earlier calls (images, kernels, clt_set_port_geometry, clt_set_tasks_geometry).  Every call on a
 * GPU shares the array: calls that overlap on one GPU (other threads, other handles) race. */
#define CLT_NOALLOC_MAX_TASKS     80000
#define CLT_NOALLOC_SCRATCH_BYTES (22u << 20)
int  clt_exec_geometry_adjoint_noalloc(clt_handle* h, double* grad_extrinsic, float* grad_disparity, double* loss);

/* Levenberg-Marquardt over the extrinsics (and the task disparities) minimising the summed loss
 * (the role of the reference's extrinsic LMA, AlignmentCorrection.java:3346-3452): gradient from
 * the adjoint; Gauss-Newton matrix from the port-centre derivatives weighted by the per-pair
 * curvatures; the disparities are eliminated per task (Schur complement) so the host solves one
 * 13 x 13 system in double precision per step. */
typedef struct clt_lma_params {
    double  lambda;             /* initial damping 0.1           (IDP lma_lambda_initial)        */
    double  lambda_scale_good;  /* 0.5                           (IDP lma_lambda_scale_good)     */
    double  lambda_scale_bad;   /* 8.0                           (IDP lma_lambda_scale_bad)      */
    double  lambda_max;         /* 100                           (IDP lma_lambda_max)            */
    double  rms_diff;           /* stop below this relative loss improvement                     */
    int32_t num_iter;           /* iterations (coarse steps included)                            */
    int32_t adjust_disparity;   /* 1: the task disparities are parameters too                    */
    /* Coarse stage (the reference's extrinsic path measures correlation-maximum offsets,
     * clt_mismatch ID:2342-2420, and solves for the extrinsics from them, AlignmentCorrection
     * AC:2891-3037, 3291-3470): the first coarse_iter iterations run convert + correlate2D, take each
     * pair's peak offset (centre of mass within coarse_radius of its maximum, +-7 px) and make the
     * Gauss-Newton step that cancels the offsets in the least-squares sense (weights: the peak
     * value less coarse_strength_floor, pairs at or below the floor left out, as the reference
     * weights its tiles, AC:1521-1525), the disparities eliminated per task; the loss-adjoint LMA
     * follows.  0: LMA only. */
    int32_t coarse_iter;        /* 0                                                             */
    int32_t reserved;
    double  coarse_radius;      /* 3.9 (max_corr_radius, ECP:2566)                               */
    /* 0.3: the correlation maximum of a pair whose offset lies beyond the +-7 px window is a
     * noise peak (about 0.23 on average against 0.5 for aligned textured tiles) */
    double  coarse_strength_floor;
} clt_lma_params;
void clt_default_lma_params(clt_lma_params* p);
/* Runs from the handle's port geometry and task disparities and leaves the handle at the solution
 * (geometry, disparities, task records).  extrinsic_out [13], disparity_out [num_tasks],
 * history [num_iter + 1] losses (initial, then after each iteration, coarse ones included); any may
 * be NULL. */
int  clt_lma_extrinsics(clt_handle* h, const clt_lma_params* p, double* extrinsic_out, float* disparity_out,
                        double* history, int32_t* iterations);

/* ---- batched frames (no reference counterpart) -------------------------------------------------
 * One handle holds num_frames frames (camera sets of the same geometry, sharing the kernels and
 * parameters), each with its own images and uneven task list.  clt_exec_batch runs the selected
 * stages for every task of every frame as one hipGraph launch (captured on first use, again when
 * the task lists change); launch shapes come from clt_autotune.  The per-frame clt_exec_* and
 * clt_set_tasks calls are not available on a batch handle. */
int  clt_create_batch(const clt_geom* geom, int32_t num_frames, clt_handle** out);
int  clt_set_image_frame(clt_handle* h, int32_t frame, int32_t cam, const float* image, int32_t stride);
int  clt_set_tasks_frame(clt_handle* h, int32_t frame, const void* tasks, int32_t num_tasks);
/* stages: 1 convert (no LPF), 2 correlate2D, 4 textures, 8 imclt_rbg with the corr_sigma LPF. */
int  clt_exec_batch(clt_handle* h, int32_t stages);
/* The batch runs as a pipeline over chunks of whole frames: each chunk's convert, textures and
 * correlate2D run in order while the previous chunks' imclt_rbg and per-tile LMA run on their own
 * streams.  Results do not depend on the chunk count. */
/* Times the candidate launch shapes on the current batch and keeps the fastest: tiles per
 * workgroup of correlate2D (1, 2, 4: its LDS per workgroup), tile rows per imclt segment (8, 16,
 * 32), tile rows per segment of the one-pass textures/imclt (4, 8, 16) with its LDS tile layout
 * (rows 20 floats apart, or 16 with swizzled 16-byte column groups, with which its even-column
 * launch runs 5 wavefronts per SIMD instead of 4) and frame chunks of the pipeline (1, 2, 4, 8, at
 * most num_frames). */
int  clt_autotune(clt_handle* h);
int  clt_get_tune(clt_handle* h, int32_t* corr_tiles_per_block, int32_t* imclt_rows);
/* values[0..n-1] (n <= 5): correlate2D tiles per workgroup, imclt rows, frame chunks, one-pass rows,
 * one-pass LDS tile layout */
int  clt_get_tune_all(clt_handle* h, int32_t* values, int32_t n);
/* Sets the first n (<= 5) of the same knobs, e.g. to reuse a clt_autotune result on another handle
 * or to pin the per-frame path's launch shapes: correlate2D tiles per workgroup 1, 2 or 4; rows
 * 1..64; chunks 1..64; layout 0 or 1.  No reference counterpart (the reference's launch shapes are
 * fixed). */
int  clt_set_tune_all(clt_handle* h, const int32_t* values, int32_t n);
/* chunks 1..64 (used as min(chunks, num_frames)); 1 until clt_autotune or this call. */
int  clt_set_batch_chunks(clt_handle* h, int32_t chunks);
int  clt_get_batch_chunks(clt_handle* h, int32_t* chunks);
/* Results of one frame, laid out as the per-frame getters (tasks in the frame's own order). */
int  clt_get_corr2d_frame(clt_handle* h, int32_t frame, float* corr, float* combo, float* disparity);
int  clt_get_textures_frame(clt_handle* h, int32_t frame, float* tiles, float* stats);
int  clt_get_rbg_frame(clt_handle* h, int32_t frame, int32_t cam, float* out);

/* ---- execution control and timing (no reference counterpart) --------------------------------- */
int  clt_set_async(clt_handle* h, int32_t async);     /* 1: exec_* return after launch        */
int  clt_sync(clt_handle* h);
/* record HIP events around each kernel; enable > 1 also creates the events of `enable` launches of
 * each kernel now (otherwise they are created at first use, inside the caller's timed region) */
int  clt_set_timing(clt_handle* h, int32_t enable);
/* Accumulated device time (ms) and launch count of a kernel since the last reset:
 * which = 0 convert, 1 imclt, 2 corr2d, 3 textures, 4 combine_textures, 5 batch graph,
 * 6 correlation adjoint, 7 geometry adjoint, 8 per-tile correlation LMA, 9 kernel preparation
 * (clt_prepare_kernels: one count per band of kernel rows, plus one for the centre slopes),
 * 10 the whole of clt_exec_corr2d_rig, 11 the one-pass textures + imclt of
 * clt_exec_textures_imclt (both column launches), 12 the fused kernel of clt_exec_convert_corr2d. */
int  clt_get_timing(clt_handle* h, int32_t which, double* total_ms, int32_t* count);
int  clt_reset_timing(clt_handle* h);

#ifdef __cplusplus
}
#endif
#endif /* CLT_TILE_PROCESSOR_H */
