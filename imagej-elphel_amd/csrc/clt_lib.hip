// clt_lib.hip -- C ABI of the MI355X CLT tile processor (include/clt_tile_processor.h).
//
// Host side of the drop-in for the reference's GPUTileProcessor (GPUTileProcessor.java): device
// buffers owned by a handle, H2D/D2H copies, kernel launches on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/clt_tile_processor.h"
#include "clt_kernels.hpp"
#include "clt_corr_lma.hpp"
#include "clt_kernel_prep.hpp"
#include "clt_rig.hpp"

namespace clt {
__constant__ Tables c_tab;
}
using namespace clt;

// ------------------------------------------------------------------------ static scratch
// The buffers of clt_exec_geometry_adjoint_noalloc, in one file-scope device array (each GPU holds
// its own copy until the process ends): per-task loss adjoint records, per-task disparity state,
// per-block partial normal systems, their column sums and the per-task disparity gradient.  Host
// code never uses the array's name as a device address: scratch_ptr() is its address on a device,
// taken once per device by hipGetSymbolAddress when the first handle is created there.
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr int NA_BLOCKS = (CLT_NOALLOC_MAX_TASKS + GEO_ADJ_TASKS - 1) / GEO_ADJ_TASKS;
constexpr size_t NA_OFF_ADJ = 0;
constexpr size_t NA_OFF_TSTATE = NA_OFF_ADJ + align256((size_t)CLT_NOALLOC_MAX_TASKS * CORR_ADJ_STRIDE * sizeof(float));
constexpr size_t NA_OFF_PART = NA_OFF_TSTATE + align256((size_t)CLT_NOALLOC_MAX_TASKS * 15 * sizeof(double));
constexpr size_t NA_OFF_SUM = NA_OFF_PART + align256((size_t)NA_BLOCKS * GEO_ADJ_SUMS * sizeof(double));
constexpr size_t NA_OFF_GD = NA_OFF_SUM + align256((size_t)GEO_ADJ_SUMS * sizeof(double));
constexpr size_t NA_BYTES = NA_OFF_GD + align256((size_t)CLT_NOALLOC_MAX_TASKS * sizeof(float));
static_assert(NA_BYTES <= CLT_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds its declared size");
static __device__ __align__(256) unsigned char scratch[CLT_NOALLOC_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// Address of the scratch array on the current device (which must be `device`), taken once.
static hipError_t scratch_init(int device, void** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

// Disparity gradient of each task (column 0 of the per-task state) as float.
__global__ __launch_bounds__(256) void k_take_col(const double* __restrict__ a, int stride, int n, float* __restrict__ out) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) out[t] = (float)a[(size_t)t * stride];
}

// ------------------------------------------------------------------------ host math helpers
// Closed-form tables (DttRad2.java:671-694, 952-1012; the DCT-IV twiddles of its FFT form)
static void make_tables(Tables* t) {
    for (int i = 0; i < 8; i++) {
        const double w = std::sin(M_PI / 16.0 * (i + 0.5));
        t->hw[i] = (float)w;
        t->hw2[i] = (float)(w * w);
        t->ihw2[i] = (float)(1.0 / (w * w));
    }
    for (int n = 0; n < 4; n++) {
        t->d4pre[n][0] = (float)(0.5 * std::cos(M_PI * (n + 0.25) / 8.0));
        t->d4pre[n][1] = (float)(-0.5 * std::sin(M_PI * (n + 0.25) / 8.0));
        t->d4post[n][0] = (float)std::cos(M_PI * n / 8.0);
        t->d4post[n][1] = (float)(-std::sin(M_PI * n / 8.0));
    }
    const double sc = std::sqrt(2.0 / 8);
    for (int j = 0; j < 8; j++)
        for (int k = 0; k < 8; k++) {
            t->ciie[j][k] = (float)(sc * std::cos(j * (2 * k + 1) * M_PI / 16.0));
            t->siie[j][k] = (float)(sc * std::sin((j + 1) * (2 * k + 1) * M_PI / 16.0));
        }
    // inverse lapped transform: orthonormal DCT-IV / DST-IV (their own inverses), unfolded to 16
    // samples with the signs of DttRad2.java:392-403 and the half window
    for (int x = 0; x < 16; x++) {
        const int j = unfold_idx(x);
        const double w = std::sin(M_PI / 16.0 * (unfold_w(x) + 0.5));
        for (int k = 0; k < 8; k++) {
            const double a = M_PI / 8.0 * (j + 0.5) * (k + 0.5);
            t->iclt_u[x][k] = (float)(w * unfold_sc(x) * sc * std::cos(a));
            t->iclt_u[x][8 + k] = (float)(w * unfold_ss(x) * sc * std::sin(a));
        }
    }
}

// setCltLpf (GPUTileProcessor.java:814-848) followed by dttt_iiie * 2n (GTP:788-798), in double.
static void lpf_coeffs(double sigma, float out[64]) {
    double fd[64];
    if (sigma == 0.0) {
        for (int i = 0; i < 64; i++) fd[i] = (i == 0) ? 1.0 : 0.0;
    } else {
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 8; j++) fd[i * 8 + j] = std::exp(-(i * i + j * j) / (2 * sigma));
    }
    double sum = 0;
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            double d = fd[i * 8 + j] * std::cos(M_PI * i / 16.0) * std::cos(M_PI * j / 16.0);
            if (i > 0) d *= 2.0;
            if (j > 0) d *= 2.0;
            sum += d;
        }
    for (int i = 0; i < 64; i++) fd[i] /= sum;
    // DCT-IIIe matrix, e_j = 0.5 for j == 0 and j == 7 (DttRad2.java:970-991)
    double m[8][8];
    const double sc = std::sqrt(2.0 / 8);
    for (int k = 0; k < 8; k++)
        for (int j = 0; j < 8; j++) m[k][j] = sc * ((j == 0 || j == 7) ? 0.5 : 1.0) * std::cos(j * (2 * k + 1) * M_PI / 16.0);
    double tmp[64], res[64];
    for (int i = 0; i < 8; i++)
        for (int k = 0; k < 8; k++) {
            double s = 0;
            for (int j = 0; j < 8; j++) s += m[k][j] * fd[i * 8 + j];
            tmp[k * 8 + i] = s;
        }
    for (int i = 0; i < 8; i++)
        for (int k = 0; k < 8; k++) {
            double s = 0;
            for (int j = 0; j < 8; j++) s += m[k][j] * tmp[i * 8 + j];
            res[i * 8 + k] = s;
        }
    for (int i = 0; i < 64; i++) out[i] = (float)(res[i] * 16.0);
}

// halfFlatTopWindow (Correlation2d.java:1047-1081)
static void half_flat_top(int n, double hwidth, double blur, bool normalize, bool notch, double scale, float* w) {
    std::vector<double> d(n);
    for (int i = 0; i < n; i++) {
        if (blur <= 0) d[i] = (i < hwidth) ? 1.0 : 0.0;
        else if (i < hwidth - blur / 2) d[i] = 1.0;
        else if (i > hwidth + blur / 2) d[i] = 0.0;
        else d[i] = 0.5 * (1.0 - std::sin(M_PI * (i - hwidth) / blur));
        if (notch) d[i] = 1.0 - d[i];
    }
    if (normalize) {
        double s = 0;
        for (int i = 0; i < n; i++) s += ((i == 0) ? 1.0 : 2.0) * d[i];
        for (int i = 0; i < n; i++) d[i] /= s;
    }
    for (int i = 0; i < n; i++) w[i] = (float)(d[i] * scale);
}

// ------------------------------------------------------------------------ handle
enum { T_CONVERT = 0, T_IMCLT, T_CORR, T_TEX, T_COMBINE, T_BATCH, T_ADJ, T_GEOADJ, T_CORR_LMA, T_KPREP, T_RIG, T_TEX_IMCLT, T_CONV_CORR, T_NUM };

struct clt_handle {
    clt_geom g{};
    GeomDev gd{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool async = false, timing = false;
    int tiles_x = 0, tiles_y = 0, num_tiles = 0, task_words = 0, max_tasks = 0;
    int num_frames = 1;            // frames (camera sets) held at once; 1 for the per-frame handle
    int cap_tasks = 0;             // num_frames * max_tasks
    long long img_fstride = 0, clt_fstride = 0;   // floats per frame of d_images / d_clt
    // device buffers
    float* d_images = nullptr;     // [C][H][W]
    float* d_kernels = nullptr;    // [C][KY][KX][K][4][h][v]
    float* d_koffs = nullptr;      // [C][KY][KX][K][8]
    int32_t* d_tasks = nullptr;    // [max_tasks][task_words]
    int32_t* d_tile_map = nullptr; // [tiles] task index or -1
    float* d_clt = nullptr;        // [C][tiles][K][4][h][v]
    float* d_lpf_rbg = nullptr;    // [3][64]
    float* d_lpf_corr = nullptr;   // [64]
    float* d_rbg = nullptr;        // [C][K][H+8][W+8]
    float* d_corr = nullptr;       // [max_tasks][P][225]
    float* d_combo = nullptr;      // [max_tasks][strip_height*15]
    float* d_disp = nullptr;       // [max_tasks][8]
    float* d_tex = nullptr;        // [max_tasks][planes][256]
    float* d_tex_stats = nullptr;  // [max_tasks][4C]
    float* d_tex_img = nullptr;    // [planes][H][W]
    float* d_extra = nullptr;      // [max_tasks]
    PortGeomDev* d_port_geom = nullptr;
    double* d_rtab = nullptr;      // reverse distortion table
    int32_t* d_gwords = nullptr;   // [max_tasks] task words of clt_set_tasks_geometry
    int32_t* d_gtxy = nullptr;     // [max_tasks]
    float* d_gdisp = nullptr;      // [max_tasks]
    double* d_gderiv = nullptr;    // [max_tasks][8][14]
    float* d_adj = nullptr;        // [cap_tasks][CORR_ADJ_STRIDE] correlation-peak loss and gradient
    float* d_offs = nullptr;       // [cap_tasks][6][3] correlation-peak offsets (clt_get_corr_offsets)
    bool adj_done = false;
    clt_port_geometry pgeom{};     // host copy of the current port geometry
    double* d_tstate = nullptr;    // [max_tasks][15] per-task LMA state
    double* d_gpart = nullptr;     // [blocks][GEO_ADJ_SUMS]
    double* d_gsum = nullptr;      // [GEO_ADJ_SUMS + 13] sums, then the extrinsic step
    float* d_gdisp2 = nullptr;     // [max_tasks] candidate disparities
    bool port_geom_set = false, tasks_from_geom = false;
    bool rig_geom = false;         // d_port_geom holds the main and the auxiliary quad of a rig
    double rig_dscale = 1.0;       // auxiliary disparity per main disparity (R_aux / R_main)
    // batched frames (clt_create_batch): host task lists per frame, flat device list with frames
    int32_t* d_task_frame = nullptr;               // [cap_tasks]
    int32_t* d_staged = nullptr;                   // [frames][max_tasks][task_words] as set per frame
    int32_t* d_fcounts = nullptr;                  // [frames] tasks per frame, then [frames] first task
    std::vector<int> frame_count;                  // tasks of each frame
    std::vector<int> frame_start;                  // first task of each frame in the flat list, + total
    std::vector<char> frame_fusable;               // every task of the frame has its own tile and non-zero flags
    bool batch_fused = false;                      // the batch runs textures + imclt_rbg as one pass
    int graph_tasks = -1;                          // task count the graph was captured for
    bool batch_dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_stages = 0;
    // launch shapes chosen by clt_autotune
    int tune_corr_tpb = 4, tune_imclt_rows = 16;
    int tune_chunks = 1;           // frame chunks of the batch pipeline
    int tex_imclt_rows = 8;        // output block rows per segment of k_tex_imclt
    int tex_lds_layout = 0;        // LDS tile layout of k_tex_imclt (TexLayout: 0 padded rows, 1 swizzled groups)
    bool tuned = false;
    bool corr_quad = false;        // correlate2D runs the quad kernel
    int corr_num_strip = 0;
    clt_corr_lma_params lp{};      // per-tile LMA chain parameters
    LmaArgs lma_args{};            // ... as launched (built by prepare_corr)
    LmaArgs* d_lma_args = nullptr; // device copy read by k_corr_lma
    LmaArgs lma_args_dev{};        // ... its contents (the upload is skipped when unchanged)
    // LMA chain scratch (fit states, fit results, per-tile values), one set per stream that runs
    // a chain: 0 the LMA stream, 1 the side stream (the rig's auxiliary quad)
    struct LmaScratch {
        LmaState* st = nullptr;    // [4 cap_tasks]
        double* res = nullptr;     // [4 cap_tasks][5]
        double* tmp = nullptr;     // [cap_tasks][8]
    } lma_scr[2];
    bool lma_args_valid = false;
    // the correlation / texture arguments last uploaded to d_corr_args / d_tex_args
    CorrArgs corr_args_dev;
    bool corr_args_valid = false;
    TexArgs tex_args_dev;
    bool tex_args_valid = false;
    bool lma_run = false;          // correlate2D is followed by k_corr_lma
    // k_corr_lma runs on its own stream, concurrently with the kernels that follow correlate2D
    // (textures, imclt, the next convert); ev_lma joins it back before its outputs are read or
    // correlate2D overwrites its inputs
    hipStream_t stream_lma = nullptr;
    hipEvent_t ev_fork = nullptr, ev_lma = nullptr;
    bool lma_pending = false;
    // a second pair of correlation records: clt_exec_corr2d called while the previous LMA may still
    // read d_corr / d_disp writes the other pair (allocated on first use), so correlate2D does not
    // wait for that LMA; ev_lma2 / lma2_pending belong to the LMA reading the other pair
    float* d_corr2 = nullptr;
    float* d_disp2 = nullptr;
    size_t corr_cap2 = 0;
    hipEvent_t ev_lma2 = nullptr;
    bool lma2_pending = false;
    // imclt_rbg of a single frame runs on a side stream from the point of its call, so it overlaps
    // what the caller launches next (correlate2D reads the same CLT); ev_side joins it back before
    // the CLT, tile map, LPF or image planes are rewritten or the images are read
    hipStream_t stream_side = nullptr;
    hipEvent_t ev_side_fork = nullptr, ev_side = nullptr;
    bool side_pending = false;
    // every task has non-zero flags and a tile of its own: tile_map names each task once, so
    // textures and imclt_rbg can run as one pass over the tile columns (k_tex_imclt)
    bool tex_fusable = false;
    int graph_epoch = 0;           // bumped whenever a captured launch would go stale
    int graph_epoch_cap = -1;
    CorrArgs* d_corr_args = nullptr;
    TexArgs* d_tex_args = nullptr;
    void* scratch = nullptr;       // the static scratch array on this handle's device
    // two-quad rig (clt_exec_corr2d_rig): cameras 0..3 the main quad, 4..7 the auxiliary one
    clt_rig_params rp{};
    float* d_rig_corr = nullptr;   // [2][max_tasks][6][225] pair correlations of the main / aux quad
    float* d_rig_disp = nullptr;   // [2][max_tasks][CLT_DISP_LAYERS]
    float* d_rig_mix = nullptr;    // [2][max_tasks][3][256] mean CLT of each quad
    float* d_rig_dirs = nullptr;   // [2][max_tasks][8] one-direction fits
    float* d_inter = nullptr;      // [max_tasks][225] inter-camera correlations
    float* d_inter_avg = nullptr;  // [max_tasks][225] their low-texture averages
    float* d_inter_rec = nullptr;  // [max_tasks][RIG_REC]
    float* d_inter_res = nullptr;  // [max_tasks][4]
    float* d_bimap = nullptr;      // [max_tasks][CLT_BI_LAYERS]
    float* d_rig_target = nullptr; // [max_tasks]
    float* d_rig_extra = nullptr;  // [2][max_tasks] each quad's disparity, for its textures
    float* d_rig_tex = nullptr;    // [2][max_tasks][planes][256] textures of the main / aux quad
    float* d_rig_tex_stats = nullptr;   // [2][max_tasks][16]
    TexArgs* d_rig_tex_args = nullptr;
    bool rig_tex_done = false;
    int rig_tex_planes = 0;        // planes the rig texture buffer holds
    int tex_planes_rig = 0;        // planes of the last rig textures
    RigArgs* d_rig_args = nullptr; // [2]: as set; with no_int_x0 for the low-texture pass
    LmaArgs* d_lma_rig = nullptr;
    bool rig_done = false, rig_target = false;
    size_t corr_cap = 0, combo_cap = 0, tex_cap = 0, tex_img_cap = 0;
    // state
    bool kernels_set[CLT_MAX_CAMS] = {};
    bool use_kernels = false;
    bool images_set[CLT_MAX_CAMS] = {};
    int num_tasks = 0;
    bool converted = false, corr_done = false, tex_done = false, combined = false;
    int combine_scale = 1;         // 1: overlap-added planes [H][W]; 2: the mosaic [2H][2W]
    int corr_pairs_at_exec = 0, corr_strip_h_at_exec = 0, tex_planes_at_exec = 0;
    float corr_sigma = 0.8f;
    std::vector<clt_pair> pairs;
    clt_corr_params cp{};
    clt_tex_params tp{};
    std::vector<int32_t> h_tile_map;
    // per kernel: event pairs of the launches since the last timing reset
    std::vector<hipEvent_t> ev_beg[T_NUM], ev_end[T_NUM];
    int ev_used[T_NUM] = {};
};

static int fail(clt_handle* h, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

#define HIPCHK(h, call)                                                                           \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return fail((h), CLT_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// Upload of a launch-argument struct to dst on the main stream, skipped when the device already
// holds the same bytes (an upload from pageable memory holds the stream for about 20 us, which a
// frame of unchanged parameters need not pay).  The kernels reading dst run on the main stream,
// after the upload that wrote their bytes.
template <class T>
static hipError_t upload_args(clt_handle* h, T* dst, const T& a, T& shadow, bool& valid) {
    if (valid && std::memcmp(&a, &shadow, sizeof(T)) == 0) return hipSuccess;
    valid = false;
    const hipError_t e = hipMemcpyAsync(dst, &a, sizeof(T), hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) return e;
    std::memcpy(&shadow, &a, sizeof(T));
    valid = true;
    return hipSuccess;
}

static int dalloc(clt_handle* h, void** p, size_t bytes) {
    if (*p) return CLT_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(h, CLT_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return CLT_OK;
}

// The main stream waits for the last per-tile LMA (its outputs are read or its inputs rewritten).
static void join_lma(clt_handle* h) {
    if (!h->lma_pending) return;
    (void)hipStreamWaitEvent(h->stream, h->ev_lma, 0);
    h->lma_pending = false;
}
static void join_lma2(clt_handle* h) {
    if (!h->lma2_pending) return;
    (void)hipStreamWaitEvent(h->stream, h->ev_lma2, 0);
    h->lma2_pending = false;
}
static void join_side(clt_handle* h) {
    if (!h->side_pending) return;
    (void)hipStreamWaitEvent(h->stream, h->ev_side, 0);
    h->side_pending = false;
}
static hipError_t sync_streams(clt_handle* h) {
    join_lma(h);
    join_lma2(h);
    join_side(h);
    return hipStreamSynchronize(h->stream);
}

static int finish(clt_handle* h) {
    if (h->async) return CLT_OK;
    HIPCHK(h, sync_streams(h));
    return CLT_OK;
}

static void tbegin(clt_handle* h, int which, hipStream_t st = nullptr) {
    if (!h->timing) return;
    const int u = h->ev_used[which];
    if (u == (int)h->ev_beg[which].size()) {
        // timing only (read after the stream is synchronised): no system-scope fence, which would
        // write back and invalidate the caches between the kernels it brackets
        hipEvent_t b = nullptr, e = nullptr;
        if (hipEventCreateWithFlags(&b, hipEventDisableSystemFence) != hipSuccess ||
            hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess)
            return;
        h->ev_beg[which].push_back(b);
        h->ev_end[which].push_back(e);
    }
    (void)hipEventRecord(h->ev_beg[which][u], st ? st : h->stream);
}
static void tend(clt_handle* h, int which, hipStream_t st = nullptr) {
    if (!h->timing) return;
    const int u = h->ev_used[which];
    if (u >= (int)h->ev_end[which].size()) return;
    (void)hipEventRecord(h->ev_end[which][u], st ? st : h->stream);
    h->ev_used[which] = u + 1;
}

// ------------------------------------------------------------------------ kernel launches
// n tasks of the device task list; task_frame: frame of each task (batches) or null (one frame).
// Tasks [t0, t0 + n) of the device task list; task_frame: frame of each task of the whole list
// (batches) or null (one frame).  The per-task kernels index their records from the pointers
// passed, so a range is the same launch on offset pointers.
static void launch_convert_range(clt_handle* h, int t0, int n, const int32_t* task_frame, int lpf_mask) {
    const long units = (long)n * h->g.num_cams;   // (task, camera); 4 per block
    const int blocks = (int)((units + 3) / 4);
    const double inv_kstep = 1.0 / h->g.kernel_step;   // exact: checked power of two
    hipLaunchKernelGGL(k_convert, dim3(blocks), dim3(64 * h->g.num_colors), 0, h->stream, h->d_images,
                       h->use_kernels ? h->d_kernels : nullptr, h->use_kernels ? h->d_koffs : nullptr,
                       h->d_tasks + (size_t)t0 * h->task_words, h->task_words, n, h->d_clt,
                       lpf_mask ? h->d_lpf_rbg : nullptr, lpf_mask, inv_kstep, task_frame ? task_frame + t0 : nullptr,
                       h->img_fstride, h->clt_fstride, h->gd);
}
static void launch_convert(clt_handle* h, int n, const int32_t* task_frame, int lpf_mask) {
    join_side(h);   // a running imclt_rbg still reads the CLT
    launch_convert_range(h, 0, n, task_frame, lpf_mask);
}

// Frames [f0, f0 + frames) of the CLT, tile maps and image planes.
static void launch_imclt(clt_handle* h, int frames, int apply_lpf, hipStream_t st = nullptr, int f0 = 0) {
    const int rows = h->tune_imclt_rows;
    const int planes = h->g.num_cams * h->g.num_colors;
    dim3 grid((h->tiles_x + 1 + IMCLT_S - 1) / IMCLT_S, (h->tiles_y + 1 + rows - 1) / rows, frames * planes);
    const size_t plane_px = (size_t)(h->g.width + 8) * (h->g.height + 8);
    hipLaunchKernelGGL(k_imclt, grid, dim3(256), 0, st ? st : h->stream, h->d_clt + (size_t)f0 * h->clt_fstride,
                       h->d_tile_map + (size_t)f0 * h->num_tiles, apply_lpf ? h->d_lpf_corr : nullptr,
                       h->d_rbg + (size_t)f0 * planes * plane_px, rows, h->clt_fstride, h->gd);
}

static void launch_corr_range(clt_handle* h, int t0, int n, const int32_t* task_frame) {
    const size_t P = h->corr_pairs_at_exec, sh = h->corr_strip_h_at_exec;
    const int32_t* tasks = h->d_tasks + (size_t)t0 * h->task_words;
    const int32_t* tf = task_frame ? task_frame + t0 : nullptr;
    float* corr = h->d_corr + (size_t)t0 * P * CLT_CORR_LEN;
    float* combo = h->d_combo ? h->d_combo + (size_t)t0 * sh * 15 : nullptr;
    float* disp = h->d_disp + (size_t)t0 * CLT_DISP_LAYERS;
    if (h->corr_quad) {
        const int tpb = h->tune_corr_tpb;
        auto kern = (h->cp.fat_zero != 0.f) ? k_corr2d_quad<true> : k_corr2d_quad<false>;
        hipLaunchKernelGGL(kern, dim3((n + tpb - 1) / tpb), dim3(64 * tpb), (size_t)tpb * CQ_WAVE_FLOATS * 4,
                           h->stream, h->d_clt, tasks, h->task_words, n, h->d_lpf_corr, h->d_corr_args, corr, combo,
                           disp, tf, h->clt_fstride, h->gd, nullptr);
    } else {
        const size_t lds = (size_t)4 * ((corr_wave_floats(h->corr_num_strip) + 3) & ~3) * 4;
        hipLaunchKernelGGL(k_corr2d<false>, dim3((n + 3) / 4), dim3(256), lds, h->stream, h->d_clt, tasks, h->task_words,
                           n, h->d_lpf_corr, h->d_corr_args, corr, combo, disp, tf, h->clt_fstride, h->gd);
    }
}
static void launch_corr(clt_handle* h, int n, const int32_t* task_frame) {
    join_lma(h);   // the previous LMA still reads d_corr and d_disp
    launch_corr_range(h, 0, n, task_frame);
}
// convert and correlate2D of tasks [t0, t0 + n) in one kernel (h->corr_quad only: one task per
// 3-wavefront block)
static void launch_convert_corr_range(clt_handle* h, int t0, int n, const int32_t* task_frame, int lpf_mask) {
    const size_t P = h->corr_pairs_at_exec, sh = h->corr_strip_h_at_exec;
    const double inv_kstep = 1.0 / h->g.kernel_step;   // exact: checked power of two
    auto kern = (h->cp.fat_zero != 0.f) ? k_convert_corr<true> : k_convert_corr<false>;
    hipLaunchKernelGGL(kern, dim3(n), dim3(192), 0, h->stream, h->d_images, h->use_kernels ? h->d_kernels : nullptr,
                       h->use_kernels ? h->d_koffs : nullptr, h->d_tasks + (size_t)t0 * h->task_words, h->task_words, n,
                       h->d_clt, lpf_mask ? h->d_lpf_rbg : nullptr, lpf_mask, inv_kstep,
                       task_frame ? task_frame + t0 : nullptr, h->img_fstride, h->clt_fstride, h->gd, h->d_lpf_corr,
                       h->d_corr_args, h->d_corr + (size_t)t0 * P * CLT_CORR_LEN,
                       h->d_combo ? h->d_combo + (size_t)t0 * sh * 15 : nullptr, h->d_disp + (size_t)t0 * CLT_DISP_LAYERS);
}

// Per-tile LMA chain over the records of tasks [t0, t0 + n) just written by correlate2D (16 tiles
// per 256-thread block), forked onto its own stream; the caller joins it (join_lma, or the end of
// a captured batch).
// The per-tile LMA chain (five kernels, clt_corr_lma.hpp) over n tasks on stream st with scratch
// set `which` (allocated by ensure_lma_scratch before; nothing here allocates, so graph capture
// works).
template <int NC>
static void lma_chain_nc(clt_handle* h, hipStream_t st, int which, const float* corr, float* disp, int n,
                         const LmaArgs* args, float* dir_out) {
    const clt_handle::LmaScratch& s = h->lma_scr[which];
    const dim3 g_tiles((n + LMA_SLOTS - 1) / LMA_SLOTS), g_one((n + 255) / 256), g_four((4 * n + 255) / 256);
    hipLaunchKernelGGL(k_lma_setup_all<NC>, g_tiles, dim3(64 * LMA_WAVES), 0, st, corr, disp, n, args, s.st);
    hipLaunchKernelGGL(k_lma_iterate<NC>, g_one, dim3(256), 0, st, s.st, n, args, s.res);
    hipLaunchKernelGGL(k_lma_setup_dirs<NC>, g_tiles, dim3(64 * LMA_WAVES), 0, st, corr, disp, n, args, s.res, s.st,
                       s.tmp, dir_out ? 1 : 0);
    hipLaunchKernelGGL(k_lma_iterate<NC>, g_four, dim3(256), 0, st, s.st, 4 * n, args, s.res);
    hipLaunchKernelGGL(k_lma_finish, g_one, dim3(256), 0, st, disp, n, args, s.tmp, s.res, dir_out);
}
static void launch_lma_chain(clt_handle* h, hipStream_t st, int which, int nc, const float* corr, float* disp, int n,
                             const LmaArgs* args, float* dir_out) {
    if (n <= 0) return;
    if (nc == 4) lma_chain_nc<4>(h, st, which, corr, disp, n, args, dir_out);
    else lma_chain_nc<3>(h, st, which, corr, disp, n, args, dir_out);
}
static int ensure_lma_scratch(clt_handle* h, int which) {
    clt_handle::LmaScratch& s = h->lma_scr[which];
    const size_t cap = h->cap_tasks;
    int rc;
    if ((rc = dalloc(h, (void**)&s.st, 4 * cap * sizeof(LmaState))) || (rc = dalloc(h, (void**)&s.res, 4 * cap * 5 * 8)) ||
        (rc = dalloc(h, (void**)&s.tmp, cap * 8 * 8)))
        return rc;
    return CLT_OK;
}

static void launch_corr_lma(clt_handle* h, int n, bool timed, int t0 = 0) {
    if (!h->lma_run) return;
    (void)hipEventRecord(h->ev_fork, h->stream);
    (void)hipStreamWaitEvent(h->stream_lma, h->ev_fork, 0);
    if (timed) tbegin(h, T_CORR_LMA, h->stream_lma);
    launch_lma_chain(h, h->stream_lma, 0, h->lma_args.adj_wyd ? 4 : 3,
                     h->d_corr + (size_t)t0 * h->corr_pairs_at_exec * CLT_CORR_LEN, h->d_disp + (size_t)t0 * CLT_DISP_LAYERS,
                     n, h->d_lma_args, nullptr);
    if (timed) tend(h, T_CORR_LMA, h->stream_lma);
    (void)hipEventRecord(h->ev_lma, h->stream_lma);
    h->lma_pending = true;
}

template <int NC>
static void launch_tex_nc(clt_handle* h, int n, const int32_t* task_frame, const float* extra, int t0) {
    const size_t planes = h->tex_planes_at_exec;
    hipLaunchKernelGGL(k_textures<NC>, dim3(n), dim3(256), 0, h->stream, h->d_clt, h->d_tasks + (size_t)t0 * h->task_words,
                       h->task_words, n, h->d_lpf_corr, extra ? extra + t0 : nullptr, h->d_tex_args,
                       h->d_tex + (size_t)t0 * planes * 256, h->d_tex_stats + (size_t)t0 * 4 * h->g.num_cams,
                       task_frame ? task_frame + t0 : nullptr, h->clt_fstride, h->gd);
}

// Frames [f0, f0 + frames) (one frame: f0 = 0, frames = 1): both column launches, in the LDS tile
// layout h->tex_lds_layout.
template <int NC, int LAY>
static void launch_tex_imclt_nc(clt_handle* h, int f0, int frames) {
    const int rows = h->tex_imclt_rows;
    const dim3 grid((h->tiles_x + 1) / 2, (h->tiles_y + 1 + rows - 1) / rows, frames);
    const size_t plane_px = (size_t)(h->g.width + 8) * (h->g.height + 8);
    // even columns store, odd columns add (with the swizzled 16-float LDS rows the even launch, which
    // reads no image planes, fits 96 VGPRs: 5 waves per SIMD for NC <= 4)
    hipLaunchKernelGGL((k_tex_imclt<NC, LAY, 0>), grid, dim3(256), 0, h->stream, h->d_clt + (size_t)f0 * h->clt_fstride,
                       h->d_tasks, h->task_words, h->d_tile_map + (size_t)f0 * h->num_tiles, h->d_lpf_corr,
                       h->d_tex_args, h->d_tex, h->d_tex_stats, h->d_rbg + (size_t)f0 * NC * 3 * plane_px, rows, h->gd,
                       h->clt_fstride);
    hipLaunchKernelGGL((k_tex_imclt<NC, LAY, 1>), grid, dim3(256), 0, h->stream, h->d_clt + (size_t)f0 * h->clt_fstride,
                       h->d_tasks, h->task_words, h->d_tile_map + (size_t)f0 * h->num_tiles, h->d_lpf_corr,
                       h->d_tex_args, h->d_tex, h->d_tex_stats, h->d_rbg + (size_t)f0 * NC * 3 * plane_px, rows, h->gd,
                       h->clt_fstride);
}
template <int NC>
static void launch_tex_imclt_lay(clt_handle* h, int f0, int frames) {
    if (h->tex_lds_layout) launch_tex_imclt_nc<NC, 1>(h, f0, frames);
    else launch_tex_imclt_nc<NC, 0>(h, f0, frames);
}
static void launch_tex_imclt(clt_handle* h, int f0 = 0, int frames = 1) {
    switch (h->g.num_cams) {
    case 2: launch_tex_imclt_lay<2>(h, f0, frames); break;
    case 3: launch_tex_imclt_lay<3>(h, f0, frames); break;
    case 4: launch_tex_imclt_lay<4>(h, f0, frames); break;
    default: launch_tex_imclt_lay<5>(h, f0, frames); break;
    }
}

static void launch_tex(clt_handle* h, int n, const int32_t* task_frame, const float* extra, int t0 = 0) {
    switch (h->g.num_cams) {
    case 2: launch_tex_nc<2>(h, n, task_frame, extra, t0); break;
    case 3: launch_tex_nc<3>(h, n, task_frame, extra, t0); break;
    case 4: launch_tex_nc<4>(h, n, task_frame, extra, t0); break;
    case 5: launch_tex_nc<5>(h, n, task_frame, extra, t0); break;
    case 6: launch_tex_nc<6>(h, n, task_frame, extra, t0); break;
    case 7: launch_tex_nc<7>(h, n, task_frame, extra, t0); break;
    default: launch_tex_nc<8>(h, n, task_frame, extra, t0); break;
    }
}

extern "C" {

int clt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void clt_default_corr_params(clt_corr_params* p) {
    p->fat_zero = 0.0f;
    p->corr_red = 0.5f;
    p->corr_blue = 0.2f;
    p->corr_offset = 0.1f;
    p->twice_diagonal = 1;
    p->min_corr = 0.02f;
    p->strip_height = 9;
    p->enhortho_width = 2.0f;
    p->enhortho_scale = 0.0f;
}

void clt_default_corr_lma_params(clt_corr_lma_params* p) {   // ImageDttParameters IDP:32-110
    std::memset(p, 0, sizeof(*p));
    p->enable = 1;
    p->mix_corr_poly = 1;
    p->min_poly_strength = 0.2;
    p->ortho_vasw_pwr = 2.0;
    p->cnvx_hwnd_size = 4;
    p->cnvx_add3x3 = 1;
    p->cnvx_weight = 0.5;
    p->lma_adjust_wm = 1;
    p->lma_adjust_wy = 0;
    p->lma_adjust_wxy = 1;
    p->lma_adjust_ag = 1;
    p->lma_half_width = 2.0;
    p->lma_cost_wy = 0.1;
    p->lma_cost_wxy = 0.1;
    p->lma_lambda_initial = 0.1;
    p->lma_lambda_scale_good = 0.5;
    p->lma_lambda_scale_bad = 8.0;
    p->lma_lambda_max = 100.0;
    p->lma_rms_diff = 0.001;
    p->lma_num_iter = 20;
    p->fo_correct = 1;
    p->fo_far = 0;
    p->fo_ortho = 1;
    p->fo_min_strength = 0.1;
    p->fo_min_eff = 0.25;
    p->fo_min_eff_ratio = 0.65;
    p->fo_max_hwidth = 3.0;
    p->fo_min_diff = 0.2;
    p->fo_overcorrection = 0.0;
    p->fo_lim_overcorr = 2.0;
}

void clt_default_tex_params(clt_tex_params* p, int32_t num_cams) {
    std::memset(p, 0, sizeof(*p));
    p->min_shot = 10.0f;
    p->scale_shot = 3.0f;
    p->diff_sigma = 5.0f;
    p->dust_remove = 1;
    p->corr_magic_scale = 0.85f;
    p->keep_weights = 1;
    if (num_cams == 4) {   // ImageDtt.java:1648-1652
        const float po[4][2] = {{-0.5f, -0.5f}, {0.5f, -0.5f}, {-0.5f, 0.5f}, {0.5f, 0.5f}};
        std::memcpy(p->port_offsets, po, sizeof(po));
    } else {
        for (int i = 0; i < num_cams && i < CLT_MAX_CAMS; i++) {
            p->port_offsets[i][0] = (float)(0.5 * std::cos(2 * M_PI * i / num_cams));
            p->port_offsets[i][1] = (float)(0.5 * std::sin(2 * M_PI * i / num_cams));
        }
    }
}

int clt_default_quad_pairs(clt_pair out[6]) {
    const clt_pair p[6] = {{0, 1, 0, 1}, {2, 3, 0, 1}, {0, 2, 1, 1}, {1, 3, 1, 1}, {0, 3, 2, 1}, {2, 1, 3, 1}};
    std::memcpy(out, p, sizeof(p));
    return 6;
}

int clt_lpf_coeffs(double sigma, float out[64]) {
    if (!out || sigma < 0) return CLT_E_ARG;
    lpf_coeffs(sigma, out);
    return CLT_OK;
}

const char* clt_last_error(const clt_handle* h) { return h ? h->err.c_str() : "null handle"; }

static int create_handle(const clt_geom* geom, int frames, clt_handle** out) {
    if (!geom || !out || frames < 1) return CLT_E_ARG;
    *out = nullptr;
    const clt_geom& g = *geom;
    if (g.width <= 16 || g.height <= 16 || (g.width & 7) || (g.height & 7) || g.num_cams < 1 ||
        g.num_cams > CLT_MAX_CAMS || (g.num_colors != 1 && g.num_colors != 3) || g.red_row < 0 || g.red_row > 1 ||
        g.red_col < 0 || g.red_col > 1 || g.kernel_step <= 0 || (g.kernel_step & (g.kernel_step - 1)) || g.kernels_hor < 0 || g.kernels_vert < 0 ||
        g.width / 8 > 0xffff || g.height / 8 > 0xffff)
        return CLT_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CLT_E_NODEVICE;
    clt_handle* h = new clt_handle();
    h->g = g;
    (void)hipGetDevice(&h->device);
    h->tiles_x = g.width / 8;
    h->tiles_y = g.height / 8;
    h->num_tiles = h->tiles_x * h->tiles_y;
    h->task_words = 2 + 2 * g.num_cams;
    h->max_tasks = g.max_tasks > 0 ? g.max_tasks : h->num_tiles;
    h->num_frames = frames;
    h->cap_tasks = frames * h->max_tasks;
    h->img_fstride = (long long)g.num_cams * g.width * g.height;
    h->clt_fstride = (long long)g.num_cams * h->num_tiles * g.num_colors * 256;
    h->frame_count.assign(frames, 0);
    h->frame_start.assign(frames + 1, 0);
    h->frame_fusable.assign(frames, 1);
    h->gd = GeomDev{g.width, g.height, g.num_cams, g.num_colors, g.red_row, g.red_col,
                    g.kernels_hor, g.kernels_vert, g.kernel_step, h->tiles_x, h->tiles_y};
    clt_default_corr_params(&h->cp);
    clt_default_corr_lma_params(&h->lp);
    clt_default_rig_params(&h->rp);
    clt_default_tex_params(&h->tp, g.num_cams);
    h->pairs.resize(0);
    if (g.num_cams == 4) {
        clt_pair p[6];
        clt_default_quad_pairs(p);
        h->pairs.assign(p, p + 6);
    }
    int rc;
    if (scratch_init(h->device, &h->scratch) != hipSuccess) {
        delete h;
        return CLT_E_HIP;
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return CLT_E_HIP;
    }
    if (hipStreamCreateWithFlags(&h->stream_lma, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_lma, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_lma2, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&h->stream_side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_side_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_side, hipEventDisableTiming) != hipSuccess) {
        (void)clt_destroy(h);
        return CLT_E_HIP;
    }
    Tables tab;
    make_tables(&tab);
    const size_t F = frames;
    if ((rc = dalloc(h, (void**)&h->d_images, F * h->img_fstride * 4)) ||
        (rc = dalloc(h, (void**)&h->d_tasks, (size_t)h->cap_tasks * h->task_words * 4)) ||
        (rc = dalloc(h, (void**)&h->d_tile_map, F * h->num_tiles * 4)) ||
        (rc = dalloc(h, (void**)&h->d_clt, F * h->clt_fstride * 4)) || (rc = dalloc(h, (void**)&h->d_lpf_rbg, 3 * 64 * 4)) ||
        (rc = dalloc(h, (void**)&h->d_lpf_corr, 64 * 4)) || (rc = dalloc(h, (void**)&h->d_corr_args, sizeof(CorrArgs))) ||
        (rc = dalloc(h, (void**)&h->d_tex_args, sizeof(TexArgs))) || (rc = dalloc(h, (void**)&h->d_extra, (size_t)h->cap_tasks * 4)) ||
        (frames > 1 && ((rc = dalloc(h, (void**)&h->d_task_frame, (size_t)h->cap_tasks * 4)) ||
                        (rc = dalloc(h, (void**)&h->d_staged, (size_t)h->cap_tasks * h->task_words * 4)) ||
                        (rc = dalloc(h, (void**)&h->d_fcounts, (size_t)2 * frames * 4))))) {
        std::string e = h->err;
        clt_destroy(h);
        return rc;
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_tab), &tab, sizeof(tab)) != hipSuccess ||
        hipMemset(h->d_images, 0, F * h->img_fstride * 4) != hipSuccess) {
        clt_destroy(h);
        return CLT_E_HIP;
    }
    h->h_tile_map.assign(F * h->num_tiles, -1);
    if (hipMemcpy(h->d_tile_map, h->h_tile_map.data(), F * h->num_tiles * 4, hipMemcpyHostToDevice) != hipSuccess) {
        clt_destroy(h);
        return CLT_E_HIP;
    }
    float lpf[64];
    lpf_coeffs(h->corr_sigma, lpf);
    (void)hipMemcpy(h->d_lpf_corr, lpf, sizeof(lpf), hipMemcpyHostToDevice);
    *out = h;
    return CLT_OK;
}

int clt_create(const clt_geom* geom, clt_handle** out) { return create_handle(geom, 1, out); }

int clt_create_batch(const clt_geom* geom, int32_t num_frames, clt_handle** out) {
    return create_handle(geom, num_frames, out);
}

int clt_destroy(clt_handle* h) {
    if (!h) return CLT_E_ARG;
    void* bufs[] = {h->d_images, h->d_kernels, h->d_koffs, h->d_tasks, h->d_tile_map, h->d_clt, h->d_lpf_rbg,
                    h->d_lpf_corr, h->d_rbg, h->d_corr, h->d_combo, h->d_disp, h->d_tex, h->d_tex_stats,
                    h->d_tex_img, h->d_extra, h->d_corr_args, h->d_tex_args, h->d_port_geom, h->d_rtab,
                    h->d_gwords, h->d_gtxy, h->d_gdisp, h->d_gderiv, h->d_task_frame, h->d_staged, h->d_fcounts,
                    h->d_adj, h->d_tstate, h->d_gpart, h->d_gsum, h->d_gdisp2, h->d_lma_args, h->d_rig_corr,
                    h->d_rig_disp, h->d_rig_mix, h->d_rig_dirs, h->d_inter, h->d_inter_avg, h->d_inter_rec,
                    h->d_inter_res, h->d_bimap, h->d_rig_target, h->d_rig_args, h->d_lma_rig, h->d_rig_extra,
                    h->d_rig_tex, h->d_rig_tex_stats, h->d_rig_tex_args, h->d_corr2, h->d_disp2,
                    h->lma_scr[0].st, h->lma_scr[0].res, h->lma_scr[0].tmp, h->lma_scr[1].st, h->lma_scr[1].res,
                    h->lma_scr[1].tmp, h->d_offs};
    if (h->stream) (void)sync_streams(h);
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    for (int i = 0; i < T_NUM; i++) {
        for (hipEvent_t e : h->ev_beg[i]) (void)hipEventDestroy(e);
        for (hipEvent_t e : h->ev_end[i]) (void)hipEventDestroy(e);
    }
    if (h->stream) (void)hipStreamDestroy(h->stream);
    if (h->stream_lma) (void)hipStreamDestroy(h->stream_lma);
    if (h->ev_side_fork) (void)hipEventDestroy(h->ev_side_fork);
    if (h->ev_side) (void)hipEventDestroy(h->ev_side);
    if (h->stream_side) (void)hipStreamDestroy(h->stream_side);
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    if (h->ev_lma) (void)hipEventDestroy(h->ev_lma);
    if (h->ev_lma2) (void)hipEventDestroy(h->ev_lma2);
    delete h;
    return CLT_OK;
}

// ------------------------------------------------------------------------ inputs
int clt_set_kernels(clt_handle* h, int32_t cam, const float* kernels, const float* offsets) {
    if (!h) return CLT_E_ARG;
    const clt_geom& g = h->g;
    if (cam < 0 || cam >= g.num_cams || !kernels || !offsets)
        return fail(h, CLT_E_ARG, "clt_set_kernels: bad camera %d or null buffer", cam);
    if (g.kernels_hor <= 0 || g.kernels_vert <= 0) return fail(h, CLT_E_ARG, "clt_set_kernels: geometry has no kernel grid");
    const size_t nk = (size_t)g.kernels_vert * g.kernels_hor * g.num_colors;
    const size_t C = g.num_cams;
    int rc;
    if ((rc = dalloc(h, (void**)&h->d_kernels, C * nk * 256 * 4)) || (rc = dalloc(h, (void**)&h->d_koffs, C * nk * 8 * 4))) return rc;
    // upload layout [v][h] per 8x8 (GPUTileProcessor.java:406-408) -> device [h][v]
    std::vector<float> tr(nk * 256);
    for (size_t k = 0; k < nk; k++)
        for (int q = 0; q < 4; q++)
            for (int v = 0; v < 8; v++)
                for (int hh = 0; hh < 8; hh++) tr[k * 256 + q * 64 + hh * 8 + v] = kernels[k * 256 + q * 64 + v * 8 + hh];
    HIPCHK(h, sync_streams(h));   // a convert in flight may read the previous kernels
    HIPCHK(h, hipMemcpy(h->d_kernels + cam * nk * 256, tr.data(), nk * 256 * 4, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_koffs + cam * nk * 8, offsets, nk * 8 * 4, hipMemcpyHostToDevice));
    h->kernels_set[cam] = true;
    h->use_kernels = true;
    for (int c = 0; c < g.num_cams; c++) h->use_kernels = h->use_kernels && h->kernels_set[c];
    h->graph_epoch++;   // a captured convert holds the kernel pointers of its capture
    return CLT_OK;
}

// calculateCLTKernel (QuadCLT.java:433-668) on the device for one camera; the raw kernel stack is
// uploaded in bands of kernel rows (at most 256 MB at a time).
int clt_prepare_kernels(clt_handle* h, int32_t cam, const float* raw, int32_t ks, const float* sensor_xy,
                        int32_t norm_kern) {
    if (!h) return CLT_E_ARG;
    const clt_geom& g = h->g;
    if (cam < 0 || cam >= g.num_cams || !raw) return fail(h, CLT_E_ARG, "clt_prepare_kernels: bad camera %d or null buffer", cam);
    if (ks < 32 || ks > 256 || (ks & 1)) return fail(h, CLT_E_ARG, "clt_prepare_kernels: raw kernel size %d (even, 32..256)", ks);
    if (g.kernels_hor <= 0 || g.kernels_vert <= 0) return fail(h, CLT_E_ARG, "clt_prepare_kernels: geometry has no kernel grid");
    const int kx = g.kernels_hor, ky = g.kernels_vert, K = g.num_colors;
    const size_t nk = (size_t)ky * kx * K, C = g.num_cams, W = (size_t)kx * ks;
    int rc;
    if ((rc = dalloc(h, (void**)&h->d_kernels, C * nk * 256 * 4)) || (rc = dalloc(h, (void**)&h->d_koffs, C * nk * 8 * 4))) return rc;
    double tab[192];
    const double sc = std::sqrt(2.0 / 8);
    for (int k = 0; k < 8; k++)   // DCT-IIIe / DST-IIIe [k][j] (DttRad2.java:970-991, 1032-1052)
        for (int j = 0; j < 8; j++) {
            tab[k * 8 + j] = sc * ((j == 0 || j == 7) ? 0.5 : 1.0) * std::cos(j * (2 * k + 1) * M_PI / 16.0);
            tab[64 + k * 8 + j] = sc * ((j == 7) ? 0.5 : 1.0) * std::sin((j + 1) * (2 * k + 1) * M_PI / 16.0);
        }
    for (int i = 0; i < 8; i++)   // norm_sym_weights (QuadCLT.java:466-474)
        for (int j = 0; j < 8; j++) tab[128 + i * 8 + j] = std::cos(M_PI * i / 16) * std::cos(M_PI * j / 16);
    const size_t row_bytes = (size_t)K * ks * W * 4;
    const int rows = (int)std::max<size_t>(1, std::min<size_t>(ky, ((size_t)256 << 20) / row_bytes));
    float *d_band = nullptr, *d_sxy = nullptr;
    double *d_tab = nullptr, *d_ctr = nullptr;
    auto cleanup = [&]() {
        (void)sync_streams(h);
        for (void* p : {(void*)d_band, (void*)d_sxy, (void*)d_tab, (void*)d_ctr})
            if (p) (void)hipFree(p);
    };
    if ((rc = dalloc(h, (void**)&d_band, (size_t)rows * row_bytes)) || (rc = dalloc(h, (void**)&d_tab, sizeof(tab))) ||
        (rc = dalloc(h, (void**)&d_ctr, nk * 2 * sizeof(double))) ||
        (sensor_xy && (rc = dalloc(h, (void**)&d_sxy, (size_t)ky * kx * 2 * 4)))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_tab, tab, sizeof(tab), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess && sensor_xy)
        e = hipMemcpyAsync(d_sxy, sensor_xy, (size_t)ky * kx * 2 * 4, hipMemcpyHostToDevice, h->stream);
    float* kern = h->d_kernels + cam * nk * 256;
    float* offs = h->d_koffs + cam * nk * 8;
    for (int row0 = 0; row0 < ky && e == hipSuccess; row0 += rows) {
        const int nr = std::min(rows, ky - row0);
        for (int col = 0; col < K && e == hipSuccess; col++)
            e = hipMemcpyAsync(d_band + (size_t)col * nr * ks * W, raw + ((size_t)col * ky + row0) * ks * W,
                               (size_t)nr * ks * W * 4, hipMemcpyHostToDevice, h->stream);
        if (e != hipSuccess) break;
        const KPrepArgs a = {ks, kx, ky, K, row0, nr, norm_kern != 0};
        const int tiles = K * nr * kx;
        tbegin(h, T_KPREP);
        hipLaunchKernelGGL(k_prepare_kernels, dim3((tiles + 3) / 4), dim3(256), 0, h->stream, d_band, d_tab, d_sxy, kern,
                           offs, d_ctr, a);
        tend(h, T_KPREP);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        tbegin(h, T_KPREP);
        hipLaunchKernelGGL(k_fill_coord_corr, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, h->stream, d_ctr, offs, kx, ky,
                           K, g.kernel_step);
        tend(h, T_KPREP);
        e = hipGetLastError();
    }
    cleanup();
    if (e != hipSuccess) return fail(h, CLT_E_HIP, "clt_prepare_kernels: %s", hipGetErrorString(e));
    h->kernels_set[cam] = true;
    h->use_kernels = true;
    for (int c = 0; c < g.num_cams; c++) h->use_kernels = h->use_kernels && h->kernels_set[c];
    h->graph_epoch++;   // a captured convert holds the kernel pointers of its capture
    return CLT_OK;
}

// The kernels of one camera as clt_set_kernels takes them (upload layout [v][h]).
int clt_get_kernels(clt_handle* h, int32_t cam, float* kernels, float* offsets) {
    if (!h) return CLT_E_ARG;
    const clt_geom& g = h->g;
    if (cam < 0 || cam >= g.num_cams || !kernels || !offsets) return fail(h, CLT_E_ARG, "clt_get_kernels: bad camera %d or null buffer", cam);
    if (!h->kernels_set[cam]) return fail(h, CLT_E_STATE, "clt_get_kernels: kernels of camera %d not set", cam);
    const size_t nk = (size_t)g.kernels_vert * g.kernels_hor * g.num_colors;
    HIPCHK(h, sync_streams(h));
    std::vector<float> tr(nk * 256);
    HIPCHK(h, hipMemcpy(tr.data(), h->d_kernels + cam * nk * 256, nk * 256 * 4, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(offsets, h->d_koffs + cam * nk * 8, nk * 8 * 4, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < nk; k++)
        for (int q = 0; q < 4; q++)
            for (int v = 0; v < 8; v++)
                for (int hh = 0; hh < 8; hh++) kernels[k * 256 + q * 64 + v * 8 + hh] = tr[k * 256 + q * 64 + hh * 8 + v];
    return CLT_OK;
}

// flattenCLTKernels (QuadCLT.java:670-734): per colour a (16 kx) x (17 ky) image; the tile of
// kernel (tx, ty) holds CC | SC over CS | SS in the CPU orientation (the transpose of the upload
// [v][h] blocks, GPUTileProcessor.java:406-408) and the 8 CltExtra values in its last row.
int clt_kernels_to_flat(const float* kernels, const float* offsets, int32_t kx, int32_t ky, int32_t colors, float* flat) {
    if (!kernels || !offsets || !flat || kx < 1 || ky < 1 || colors < 1) return CLT_E_ARG;
    const size_t Wf = (size_t)16 * kx, Hf = (size_t)17 * ky;
    std::memset(flat, 0, (size_t)colors * Wf * Hf * 4);
    for (int ty = 0; ty < ky; ty++)
        for (int tx = 0; tx < kx; tx++)
            for (int col = 0; col < colors; col++) {
                const size_t k = ((size_t)ty * kx + tx) * colors + col;
                float* base = flat + (size_t)col * Wf * Hf + (size_t)ty * 17 * Wf + (size_t)tx * 16;
                for (int q = 0; q < 4; q++)
                    for (int i = 0; i < 8; i++)
                        for (int j = 0; j < 8; j++)
                            base[(i + ((q & 2) ? 8 : 0)) * Wf + j + ((q & 1) ? 8 : 0)] = kernels[k * 256 + q * 64 + j * 8 + i];
                for (int i = 0; i < 8; i++) base[16 * Wf + i] = offsets[k * 8 + i];
            }
    return CLT_OK;
}

// extractCLTKernels (QuadCLT.java:790-857): the inverse of clt_kernels_to_flat.
int clt_kernels_from_flat(const float* flat, int32_t kx, int32_t ky, int32_t colors, float* kernels, float* offsets) {
    if (!kernels || !offsets || !flat || kx < 1 || ky < 1 || colors < 1) return CLT_E_ARG;
    const size_t Wf = (size_t)16 * kx, Hf = (size_t)17 * ky;
    for (int ty = 0; ty < ky; ty++)
        for (int tx = 0; tx < kx; tx++)
            for (int col = 0; col < colors; col++) {
                const size_t k = ((size_t)ty * kx + tx) * colors + col;
                const float* base = flat + (size_t)col * Wf * Hf + (size_t)ty * 17 * Wf + (size_t)tx * 16;
                for (int q = 0; q < 4; q++)
                    for (int i = 0; i < 8; i++)
                        for (int j = 0; j < 8; j++)
                            kernels[k * 256 + q * 64 + j * 8 + i] = base[(i + ((q & 2) ? 8 : 0)) * Wf + j + ((q & 1) ? 8 : 0)];
                for (int i = 0; i < 8; i++) offsets[k * 8 + i] = base[16 * Wf + i];
            }
    return CLT_OK;
}

int clt_clear_kernels(clt_handle* h) {
    if (!h) return CLT_E_ARG;
    for (int c = 0; c < CLT_MAX_CAMS; c++) h->kernels_set[c] = false;
    h->use_kernels = false;
    h->graph_epoch++;   // a captured convert holds the kernel pointers of its capture
    return CLT_OK;
}

int clt_set_image(clt_handle* h, int32_t cam, const float* image, int32_t stride) {
    if (!h) return CLT_E_ARG;
    const clt_geom& g = h->g;
    if (cam < 0 || cam >= g.num_cams || !image || stride < g.width)
        return fail(h, CLT_E_ARG, "clt_set_image: bad camera %d, null image or stride %d < width %d", cam, stride, g.width);
    HIPCHK(h, sync_streams(h));   // the stream may still read the previous frame
    HIPCHK(h, hipMemcpy2D(h->d_images + (size_t)cam * g.width * g.height, (size_t)g.width * 4, image, (size_t)stride * 4,
                          (size_t)g.width * 4, g.height, hipMemcpyHostToDevice));
    h->images_set[cam] = true;
    return CLT_OK;
}

int clt_set_tasks(clt_handle* h, const void* tasks, int32_t num_tasks) {
    if (!h) return CLT_E_ARG;
    if (h->num_frames > 1) return fail(h, CLT_E_UNSUPPORTED, "clt_set_tasks: a batch handle takes clt_set_tasks_frame");
    if (num_tasks < 0 || num_tasks > h->max_tasks || (num_tasks > 0 && !tasks))
        return fail(h, CLT_E_ARG, "clt_set_tasks: %d tasks (capacity %d)", num_tasks, h->max_tasks);
    const int32_t* t = (const int32_t*)tasks;
    HIPCHK(h, sync_streams(h));
    std::fill(h->h_tile_map.begin(), h->h_tile_map.end(), -1);
    bool fusable = true;
    for (int i = 0; i < num_tasks; i++) {
        const int32_t* r = t + (size_t)i * h->task_words;
        const int tx = r[1] & 0xffff, ty = (r[1] >> 16) & 0xffff;
        if (tx >= h->tiles_x || ty >= h->tiles_y)
            return fail(h, CLT_E_ARG, "clt_set_tasks: task %d tile (%d,%d) outside %dx%d", i, tx, ty, h->tiles_x, h->tiles_y);
        int32_t& m = h->h_tile_map[(size_t)ty * h->tiles_x + tx];
        fusable = fusable && r[0] != 0 && m < 0;
        if (r[0] != 0) m = i;
    }
    h->tex_fusable = fusable;
    if (num_tasks > 0) HIPCHK(h, hipMemcpy(h->d_tasks, tasks, (size_t)num_tasks * h->task_words * 4, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_tile_map, h->h_tile_map.data(), (size_t)h->num_tiles * 4, hipMemcpyHostToDevice));
    h->batch_dirty = true;   // the device task list now holds frame 0 only
    h->num_tasks = num_tasks;
    h->converted = h->corr_done = h->tex_done = h->combined = h->adj_done = h->rig_done = false;
    h->tasks_from_geom = false;
    return CLT_OK;
}

// ------------------------------------------------------------------------ port geometry
void clt_default_port_geometry(clt_port_geometry* p, int32_t width, int32_t height) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->num_sensors = 4;
    p->width = width;
    p->height = height;
    p->focal_length = 4.5;         // GC:37
    p->pixel_size = 2.2;           // GC:39
    p->distortion_radius = 2.8512; // GC:38
    static const double rxy[4][2] = {{-0.5, -0.5}, {0.5, -0.5}, {-0.5, 0.5}, {0.5, 0.5}};   // GC:78
    for (int i = 0; i < 4; i++) {
        p->rxy[i][0] = rxy[i][0];
        p->rxy[i][1] = rxy[i][1];
        p->pxy0[i][0] = 0.5 * width;
        p->pxy0[i][1] = 0.5 * height;
    }
}

static void mat3(const double* a, const double* b, double* c) {
    double t[9];
    for (int r = 0; r < 3; r++)
        for (int k = 0; k < 3; k++) t[r * 3 + k] = a[r * 3] * b[k] + a[r * 3 + 1] * b[3 + k] + a[r * 3 + 2] * b[6 + k];
    std::memcpy(c, t, sizeof(t));
}

static void port_geom_dev(const clt_port_geometry* p, PortGeomDev& pg) {
    std::memset(&pg, 0, sizeof(pg));
    const double* v = p->extrinsic;
    for (int i = 0; i < 4; i++) {   // CorrVector angles (GC:1217-1269) and matrices (GC:1043-1147)
        const double tl = (i < 3) ? v[i] : -(v[0] + v[1] + v[2]);
        const double az = (i < 3) ? v[3 + i] : -(v[3] + v[4] + v[5]);
        const double rl = v[6 + i] + p->roll_deg[i] * M_PI / 180.0;
        const double zm = 1.0 + ((i < 3) ? v[10 + i] : -(v[10] + v[11] + v[12]));
        const double ca = std::cos(az), sa = std::sin(az), ct = std::cos(tl), st = std::sin(tl);
        const double cr = std::cos(rl), sr = std::sin(rl);
        // azimuth (sign -1), tilt, roll with zoom; "inverted" rotations as in the reference
        const double m_az[9] = {ca, 0, -sa, 0, 1, 0, sa, 0, ca};
        const double m_t[9] = {1, 0, 0, 0, ct, st, 0, -st, ct};
        const double m_r[9] = {cr * zm, sr * zm, 0, -sr * zm, cr * zm, 0, 0, 0, 1};
        const double d_az[9] = {-sa, 0, -ca, 0, 0, 0, ca, 0, -sa};
        const double d_t[9] = {0, 0, 0, 0, -st, ct, 0, -ct, -st};
        const double d_r[9] = {-sr * zm, cr * zm, 0, -cr * zm, -sr * zm, 0, 0, 0, 0};
        const double d_z[9] = {cr, sr, 0, -sr, cr, 0, 0, 0, 0};
        double ta[9], tmp[9];
        mat3(m_t, m_az, ta);
        mat3(m_r, ta, pg.rots[i]);
        mat3(m_t, d_az, tmp);
        mat3(m_r, tmp, pg.drots[i][0]);
        mat3(d_t, m_az, tmp);
        mat3(m_r, tmp, pg.drots[i][1]);
        mat3(d_r, ta, pg.drots[i][2]);
        mat3(d_z, ta, pg.drots[i][3]);
        pg.rxy[i][0] = p->rxy[i][0];
        pg.rxy[i][1] = p->rxy[i][1];
        pg.pxy0[i][0] = p->pxy0[i][0];
        pg.pxy0[i][1] = p->pxy0[i][1];
    }
    for (int j = 0; j < 7; j++) pg.dist[j] = p->distortion[j];
    pg.fl_pix = p->focal_length / (0.001 * p->pixel_size);
    pg.ri_scale = 0.001 * p->pixel_size / p->distortion_radius;
    pg.rd_scale = 0.001 * p->pixel_size / p->distortion_radius;
    pg.half_w = 0.5 * p->width;
    pg.half_h = 0.5 * p->height;
    pg.step_r = 0.001;
    pg.tab_len = (int)std::ceil(2.0 / pg.step_r) + 1;
}

// reverse distortion table (calcReverseDistortionTable, GC:2707-2756): r / rD by Newton steps
static std::vector<double> reverse_distortion_tab(const clt_port_geometry* p) {
    const double step = 0.001, max_r = 2.0;
    const int len = (int)std::ceil(max_r / step) + 1;
    std::vector<double> tab(len, std::nan(""));
    const double C = p->distortion[0], B = p->distortion[1], A = p->distortion[2], A5 = p->distortion[3];
    const double A6 = p->distortion[4], A7 = p->distortion[5], A8 = p->distortion[6];
    const double d0 = 1.0 - A8 - A7 - A6 - A5 - A - B - C;
    auto k_of = [&](double r) { return (((((((A8 * r + A7) * r + A6) * r + A5) * r + A) * r + B) * r + C) * r + d0); };
    auto dk_of = [&](double r) {
        return (((((((8 * A8 * r + 7 * A7) * r + 6 * A6) * r + 5 * A5) * r + 4 * A) * r + 3 * B) * r + 2 * C) * r + d0);
    };
    tab[0] = 1.0 / d0;
    double rprev = 0.0, slope = 1.0;
    for (int i = 1; i < len; i++) {
        const double rd_target = step * i;
        double r = rprev + step / slope;
        for (int it = 0; it < 1000; it++) {
            slope = dk_of(r);
            const double rd = r * k_of(r);
            if (slope < 0.1) break;   // too much distortion: the rest of the table stays NaN
            if (std::fabs(rd - rd_target) < 1e-8) break;
            r += (rd_target - rd) / slope;
        }
        if (slope < 0.1) break;
        rprev = r;
        tab[i] = r / rd_target;
    }
    return tab;
}

// ns PortGeomDev records (1: a quad, 2: main and auxiliary quad) and the reverse distortion table
static int upload_port_geometry(clt_handle* h, const PortGeomDev* pg, int ns, const std::vector<double>& tab) {
    int rc;
    if ((rc = dalloc(h, (void**)&h->d_port_geom, 2 * sizeof(PortGeomDev))) ||
        (rc = dalloc(h, (void**)&h->d_rtab, sizeof(double) * tab.size())))
        return rc;
    HIPCHK(h, sync_streams(h));
    HIPCHK(h, hipMemcpy(h->d_port_geom, pg, ns * sizeof(PortGeomDev), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_rtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
    return CLT_OK;
}

static bool geometry_ok(const clt_port_geometry* p) {
    return p && p->num_sensors == 4 && p->width > 0 && p->height > 0 && p->pixel_size > 0 && p->distortion_radius > 0 &&
           p->focal_length > 0;
}

int clt_set_port_geometry(clt_handle* h, const clt_port_geometry* p) {
    if (!h) return CLT_E_ARG;
    if (!geometry_ok(p) || h->g.num_cams != 4)
        return fail(h, CLT_E_ARG, "clt_set_port_geometry: needs a quad camera handle and positive sizes");
    PortGeomDev pg;
    port_geom_dev(p, pg);
    int rc;
    if ((rc = upload_port_geometry(h, &pg, 1, reverse_distortion_tab(p)))) return rc;
    h->pgeom = *p;
    h->port_geom_set = true;
    h->rig_geom = false;
    return CLT_OK;
}

// clt_bi_quad's port centres (ID:8970-8993): the auxiliary quad undistorts the tile centre with the
// main camera's model (gc_main, GC:2179-2184), uses rXY_aux (RigOffset.recalcRXY GC:728-745) and
// its CorrVector rotations times the rig matrix (getRotMatrices(rigMatrix) GC:1016-1026,
// RigOffset.getRotMatrix GC:770-803), at disparity * R_aux / R_main (ID:8968).
int clt_set_port_geometry_rig(clt_handle* h, const clt_port_geometry* main, const clt_port_geometry* aux,
                              const clt_rig_offset* rig) {
    if (!h) return CLT_E_ARG;
    if (!geometry_ok(main) || !geometry_ok(aux) || !rig || h->g.num_cams != 8 || !(rig->main_disparity_radius > 0) ||
        !(rig->aux_disparity_radius > 0))
        return fail(h, CLT_E_ARG, "clt_set_port_geometry_rig: needs an 8-camera handle, positive sizes and radii");
    PortGeomDev pg[2];
    port_geom_dev(main, pg[0]);
    port_geom_dev(aux, pg[1]);
    const double ca = std::cos(rig->aux_azimuth), sa = std::sin(rig->aux_azimuth);
    const double ct = std::cos(rig->aux_tilt), st = std::sin(rig->aux_tilt);
    const double zoom = 1.0 + rig->aux_zoom, cr = std::cos(rig->aux_roll) * zoom, sr = std::sin(rig->aux_roll) * zoom;
    const double m_az[9] = {ca, 0, -sa, 0, 1, 0, sa, 0, ca};   // ROT_AZ_SGN = -1
    const double m_t[9] = {1, 0, 0, 0, ct, st, 0, -st, ct};
    const double m_r[9] = {cr, sr, 0, -sr, cr, 0, 0, 0, 1};
    double ta[9], rm[9];
    mat3(m_t, m_az, ta);
    mat3(m_r, ta, rm);
    const double xc = rig->baseline * std::cos(rig->aux_angle) / rig->aux_disparity_radius;
    const double yc = rig->baseline * std::sin(rig->aux_angle) / rig->aux_disparity_radius;
    const double c = std::cos(rig->aux_roll), sn = std::sin(rig->aux_roll);
    for (int i = 0; i < 4; i++) {
        double r[9];
        mat3(pg[1].rots[i], rm, r);
        std::memcpy(pg[1].rots[i], r, sizeof(r));
        pg[1].rxy[i][0] = xc + c * aux->rxy[i][0] + sn * aux->rxy[i][1];
        pg[1].rxy[i][1] = yc - sn * aux->rxy[i][0] + c * aux->rxy[i][1];
    }
    pg[1].half_w = pg[0].half_w;   // the tile centre through the main camera's distortion
    pg[1].half_h = pg[0].half_h;
    pg[1].rd_scale = pg[0].rd_scale;
    int rc;
    if ((rc = upload_port_geometry(h, pg, 2, reverse_distortion_tab(main)))) return rc;
    h->pgeom = *main;
    h->port_geom_set = true;
    h->rig_geom = true;
    h->rig_dscale = rig->aux_disparity_radius / rig->main_disparity_radius;
    return CLT_OK;
}

static int launch_port_centres(clt_handle* h, double* deriv) {
    const int n = h->num_tasks;
    if (n == 0) return CLT_OK;
    const int ns = h->rig_geom ? 8 : 4;
    hipLaunchKernelGGL(k_port_centres, dim3((ns * n + 255) / 256), dim3(256), 0, h->stream, h->d_port_geom, h->d_rtab,
                       h->d_gwords, h->d_gtxy, h->d_gdisp, n, h->task_words, h->d_tasks, deriv, h->rig_geom ? 1 : 0,
                       h->rig_dscale);
    HIPCHK(h, hipGetLastError());
    return CLT_OK;
}

int clt_set_tasks_geometry(clt_handle* h, const int32_t* task, const int32_t* txy, const float* disparity,
                           int32_t num_tasks) {
    if (!h) return CLT_E_ARG;
    if (!h->port_geom_set) return fail(h, CLT_E_STATE, "clt_set_tasks_geometry: no port geometry set");
    if (num_tasks < 0 || num_tasks > h->max_tasks || (num_tasks > 0 && (!task || !txy || !disparity)))
        return fail(h, CLT_E_ARG, "clt_set_tasks_geometry: %d tasks (capacity %d)", num_tasks, h->max_tasks);
    int rc;
    if ((rc = dalloc(h, (void**)&h->d_gwords, (size_t)h->max_tasks * 4)) ||
        (rc = dalloc(h, (void**)&h->d_gtxy, (size_t)h->max_tasks * 4)) ||
        (rc = dalloc(h, (void**)&h->d_gdisp, (size_t)h->max_tasks * 4)))
        return rc;
    HIPCHK(h, sync_streams(h));
    std::fill(h->h_tile_map.begin(), h->h_tile_map.end(), -1);
    bool fusable = true;
    for (int i = 0; i < num_tasks; i++) {
        const int tx = txy[i] & 0xffff, ty = (txy[i] >> 16) & 0xffff;
        if (tx >= h->tiles_x || ty >= h->tiles_y)
            return fail(h, CLT_E_ARG, "clt_set_tasks_geometry: task %d tile (%d,%d) outside %dx%d", i, tx, ty, h->tiles_x, h->tiles_y);
        int32_t& m = h->h_tile_map[(size_t)ty * h->tiles_x + tx];
        fusable = fusable && task[i] != 0 && m < 0;
        if (task[i] != 0) m = i;
    }
    h->tex_fusable = fusable;
    if (num_tasks > 0) {
        HIPCHK(h, hipMemcpy(h->d_gwords, task, (size_t)num_tasks * 4, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->d_gtxy, txy, (size_t)num_tasks * 4, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->d_gdisp, disparity, (size_t)num_tasks * 4, hipMemcpyHostToDevice));
    }
    HIPCHK(h, hipMemcpy(h->d_tile_map, h->h_tile_map.data(), (size_t)h->num_tiles * 4, hipMemcpyHostToDevice));
    h->num_tasks = num_tasks;
    h->converted = h->corr_done = h->tex_done = h->combined = h->adj_done = h->rig_done = false;
    h->tasks_from_geom = true;
    h->batch_dirty = true;
    if ((rc = launch_port_centres(h, nullptr))) return rc;
    return finish(h);
}

int clt_get_tasks(clt_handle* h, void* tasks) {
    if (!h || !tasks) return CLT_E_ARG;
    HIPCHK(h, sync_streams(h));
    if (h->num_tasks > 0)
        HIPCHK(h, hipMemcpy(tasks, h->d_tasks, (size_t)h->num_tasks * h->task_words * 4, hipMemcpyDeviceToHost));
    return CLT_OK;
}

int clt_get_port_derivatives(clt_handle* h, double* out) {
    if (!h || !out) return CLT_E_ARG;
    if (h->rig_geom) return fail(h, CLT_E_UNSUPPORTED, "clt_get_port_derivatives: a single quad only");
    if (!h->tasks_from_geom) return fail(h, CLT_E_STATE, "clt_get_port_derivatives: tasks were not set from the port geometry");
    int rc;
    if ((rc = dalloc(h, (void**)&h->d_gderiv, (size_t)h->max_tasks * 8 * 14 * sizeof(double)))) return rc;
    if ((rc = launch_port_centres(h, h->d_gderiv))) return rc;
    HIPCHK(h, sync_streams(h));
    if (h->num_tasks > 0)
        HIPCHK(h, hipMemcpy(out, h->d_gderiv, (size_t)h->num_tasks * 8 * 14 * sizeof(double), hipMemcpyDeviceToHost));
    return CLT_OK;
}

int clt_set_lpf_rbg(clt_handle* h, float sigma_r, float sigma_b, float sigma_g) {
    if (!h) return CLT_E_ARG;
    if (sigma_r < 0 || sigma_b < 0 || sigma_g < 0) return fail(h, CLT_E_ARG, "clt_set_lpf_rbg: negative sigma");
    float l[3][64];
    lpf_coeffs(sigma_r, l[0]);
    lpf_coeffs(sigma_b, l[1]);
    lpf_coeffs(sigma_g, l[2]);
    HIPCHK(h, sync_streams(h));   // kernels in flight may read the table
    HIPCHK(h, hipMemcpy(h->d_lpf_rbg, l, sizeof(l), hipMemcpyHostToDevice));
    return CLT_OK;
}

int clt_set_corr_sigma(clt_handle* h, float sigma) {
    if (!h) return CLT_E_ARG;
    if (sigma < 0) return fail(h, CLT_E_ARG, "clt_set_corr_sigma: negative sigma");
    float l[64];
    lpf_coeffs(sigma, l);
    HIPCHK(h, sync_streams(h));   // kernels in flight may read the table
    HIPCHK(h, hipMemcpy(h->d_lpf_corr, l, sizeof(l), hipMemcpyHostToDevice));
    h->corr_sigma = sigma;
    return CLT_OK;
}

int clt_set_pairs(clt_handle* h, const clt_pair* pairs, int32_t num_pairs) {
    if (!h) return CLT_E_ARG;
    if (num_pairs < 1 || num_pairs > CLT_MAX_PAIRS || !pairs) return fail(h, CLT_E_ARG, "clt_set_pairs: %d pairs", num_pairs);
    int nstrip = 0;
    for (int i = 0; i < num_pairs; i++) {
        const clt_pair& p = pairs[i];
        if (p.cam1 < 0 || p.cam1 >= h->g.num_cams || p.cam2 < 0 || p.cam2 >= h->g.num_cams || p.dir < -1 || p.dir > 3 ||
            (p.dir >= 0 && p.scale < 1))
            return fail(h, CLT_E_ARG, "clt_set_pairs: bad pair %d", i);
        nstrip += p.dir >= 0;
    }
    if (nstrip > CORR_MAX_STRIP_PAIRS) return fail(h, CLT_E_UNSUPPORTED, "clt_set_pairs: %d oriented pairs > %d", nstrip, CORR_MAX_STRIP_PAIRS);
    h->pairs.assign(pairs, pairs + num_pairs);
    h->corr_done = false;
    h->graph_epoch++;
    return CLT_OK;
}

int clt_set_corr_params(clt_handle* h, const clt_corr_params* p) {
    if (!h || !p) return CLT_E_ARG;
    if (p->strip_height < 1 || p->strip_height > 16) return fail(h, CLT_E_ARG, "strip_height %d", p->strip_height);
    h->cp = *p;
    h->graph_epoch++;   // the correlation kernel variant follows fat_zero
    return CLT_OK;
}

int clt_set_corr_lma_params(clt_handle* h, const clt_corr_lma_params* p) {
    if (!h || !p) return CLT_E_ARG;
    if (p->cnvx_hwnd_size < 1 || p->cnvx_hwnd_size > 4 || p->lma_num_iter < 0 || !(p->lma_lambda_initial >= 0) ||
        !(p->lma_half_width > 0))
        return fail(h, CLT_E_ARG, "clt_set_corr_lma_params: half window %d (1..4), %d iterations", p->cnvx_hwnd_size,
                    p->lma_num_iter);
    h->lp = *p;
    h->graph_epoch++;   // the LMA arguments are launch arguments
    return CLT_OK;
}

int clt_set_tex_params(clt_handle* h, const clt_tex_params* p) {
    if (!h || !p) return CLT_E_ARG;
    if (p->diff_sigma <= 0 || p->corr_magic_scale == 0 || p->min_shot <= 0) return fail(h, CLT_E_ARG, "bad texture parameters");
    h->tp = *p;
    return CLT_OK;
}

// ------------------------------------------------------------------------ kernels
int clt_exec_convert(clt_handle* h, int32_t lpf_mask) {
    if (!h) return CLT_E_ARG;
    for (int c = 0; c < h->g.num_cams; c++)
        if (!h->images_set[c]) return fail(h, CLT_E_STATE, "clt_exec_convert: image of camera %d not set", c);
    if (h->num_tasks <= 0) return fail(h, CLT_E_STATE, "clt_exec_convert: no tasks");
    if (h->num_frames > 1) return fail(h, CLT_E_UNSUPPORTED, "clt_exec_convert: a batch handle runs clt_exec_batch");
    tbegin(h, T_CONVERT);
    launch_convert(h, h->num_tasks, nullptr, lpf_mask);
    tend(h, T_CONVERT);
    HIPCHK(h, hipGetLastError());
    h->converted = true;
    h->corr_done = h->tex_done = h->combined = h->rig_done = false;
    return finish(h);
}

int clt_exec_imclt(clt_handle* h, int32_t apply_lpf) {
    if (!h) return CLT_E_ARG;
    if (h->num_frames > 1) return fail(h, CLT_E_UNSUPPORTED, "clt_exec_imclt: a batch handle runs clt_exec_batch");
    if (!h->converted) return fail(h, CLT_E_STATE, "clt_exec_imclt: no converted tiles");
    const clt_geom& g = h->g;
    int rc;
    if ((rc = dalloc(h, (void**)&h->d_rbg,
                     (size_t)h->num_frames * g.num_cams * g.num_colors * (g.width + 8) * (g.height + 8) * 4)))
        return rc;
    join_side(h);   // the previous imclt_rbg writes the same planes
    HIPCHK(h, hipEventRecord(h->ev_side_fork, h->stream));
    HIPCHK(h, hipStreamWaitEvent(h->stream_side, h->ev_side_fork, 0));
    tbegin(h, T_IMCLT, h->stream_side);
    launch_imclt(h, 1, apply_lpf, h->stream_side);
    tend(h, T_IMCLT, h->stream_side);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->ev_side, h->stream_side));
    h->side_pending = true;
    return finish(h);
}

static int ensure_capacity(clt_handle* h, float** p, size_t* cap, size_t floats) {
    if (*p && *cap >= floats) return CLT_OK;
    if (*p) {
        (void)sync_streams(h);
        (void)hipFree(*p);
        *p = nullptr;
    }
    h->graph_epoch++;   // a captured graph holds the old address
    int rc = dalloc(h, (void**)p, floats * 4);
    if (rc == CLT_OK) *cap = floats;
    return rc;
}

// LMA launch arguments for the current pairs: fit 0 over every oriented pair, fits 1..4 over the
// horizontal, vertical, main- and other-diagonal pairs (corr4dirsLMA, C2:1591-1630); each fit's pairs
// grouped by {diagonal, scale} in the order of GROUPS (C2:72-78) and combined as
// combineCompatiblePairs does (C2:462-488).
// The LMA fits of a pair list (the fit over all oriented pairs, then one per direction) and the
// parameters of lp into a.
static int fill_lma_args(clt_handle* h, const std::vector<clt_pair>& pairs, const clt_corr_lma_params& lp, LmaArgs& a) {
    std::memset(&a, 0, sizeof(a));
    static const int groups[6][2] = {{0, 1}, {1, 1}, {0, 2}, {1, 2}, {0, 4}, {1, 4}};
    const int P = (int)pairs.size();
    for (int f = 0; f < LMA_FIT_SINGLE; f++) {
        LmaFit& F = a.fit[f];
        for (int gi = 0; gi < 6; gi++) {
            const int diag = groups[gi][0], scale = groups[gi][1];
            int members[CLT_MAX_PAIRS], n = 0;
            for (int p = 0; p < P; p++) {
                const clt_pair& q = pairs[p];
                const bool sel = (f == 0) ? q.dir >= 0 : q.dir == f - 1;
                if (sel && q.dir >= 0 && ((q.dir >= 2) == (diag != 0)) && q.scale == scale) members[n++] = p;
            }
            if (n == 0) continue;
            if (F.ngroups == ((f == 0) ? LMA_MAXG : 1) || n > ((f == 0) ? LMA_GPAIRS : LMA_DPAIRS))
                return fail(h, CLT_E_UNSUPPORTED,
                            "per-tile LMA: over all pairs at most %d groups of %d pairs, per direction one group of "
                            "%d pairs (disable it with clt_set_corr_lma_params)", LMA_MAXG, LMA_GPAIRS, LMA_DPAIRS);
            LmaGroup& G = F.g[F.ngroups++];
            G.npairs = n;
            G.diag = diag;
            G.scale = scale;
            for (int i = 0; i < n; i++) {
                const int d = pairs[members[i]].dir;
                G.pair[i] = members[i];
                G.xform[i] = (d == 1) ? 1 : ((d == 3) ? 2 : 0);
            }
        }
    }
    a.min_poly_strength = lp.min_poly_strength;
    a.vasw_pwr = lp.ortho_vasw_pwr;
    a.nc_cost = lp.cnvx_weight;
    a.half_width = lp.lma_half_width;
    a.cost_wm = lp.lma_cost_wy;
    a.cost_wxy = lp.lma_cost_wxy;
    a.lambda0 = lp.lma_lambda_initial;
    a.l_good = lp.lma_lambda_scale_good;
    a.l_bad = lp.lma_lambda_scale_bad;
    a.l_max = lp.lma_lambda_max;
    a.rms_diff = lp.lma_rms_diff;
    a.fo_min_strength = lp.fo_min_strength;
    a.fo_min_eff = lp.fo_min_eff;
    a.fo_min_eff_ratio = lp.fo_min_eff_ratio;
    a.fo_max_hwidth = lp.fo_max_hwidth;
    a.fo_min_diff = lp.fo_min_diff;
    a.fo_overcorr = lp.fo_overcorrection;
    a.fo_lim_overcorr = lp.fo_lim_overcorr;
    a.hwin = lp.cnvx_hwnd_size;
    a.add3x3 = lp.cnvx_add3x3;
    a.adj_wm = lp.lma_adjust_wm;
    a.adj_wyd = lp.lma_adjust_wy;
    a.adj_wxy = lp.lma_adjust_wxy;
    a.adj_ag = lp.lma_adjust_ag;
    a.num_iter = lp.lma_num_iter;
    a.mix = lp.mix_corr_poly;
    a.fo_correct = lp.fo_correct;
    a.fo_far = lp.fo_far;
    a.fo_ortho = lp.fo_ortho;
    a.npairs = P;
    return CLT_OK;
}

static int build_lma_args(clt_handle* h, bool have_records) {
    LmaArgs& a = h->lma_args;
    std::memset(&a, 0, sizeof(a));
    h->lma_run = false;
    if (!h->lp.enable || !have_records) return CLT_OK;
    int rc;
    if ((rc = fill_lma_args(h, h->pairs, h->lp, a))) return rc;
    h->lma_run = a.fit[0].ngroups > 0;
    if ((rc = dalloc(h, (void**)&h->d_lma_args, sizeof(LmaArgs)))) return rc;
    if (h->lma_run && (rc = ensure_lma_scratch(h, 0))) return rc;
    if (h->lma_args_valid && std::memcmp(&a, &h->lma_args_dev, sizeof(a)) == 0) return CLT_OK;   // both zero-filled
    join_lma(h);    // a running LMA still reads the previous arguments
    join_lma2(h);
    HIPCHK(h, hipMemcpyAsync(h->d_lma_args, &a, sizeof(a), hipMemcpyHostToDevice, h->stream));
    std::memcpy(&h->lma_args_dev, &a, sizeof(a));
    h->lma_args_valid = true;
    return CLT_OK;
}

// CorrArgs of the current pairs and parameters (host only, no allocation); *quad: the reference
// quad configuration, which runs the specialised kernel (pairs as compile-time constants).
static int fill_corr_args(clt_handle* h, CorrArgs& a, bool* quad) {
    if (h->pairs.empty()) return fail(h, CLT_E_STATE, "clt_exec_corr2d: no correlation pairs set");
    const clt_geom& g = h->g;
    const int P = (int)h->pairs.size(), hw = h->cp.strip_height;
    std::memset(&a, 0, sizeof(a));
    a.num_pairs = P;
    int ss_min = 1 << 30;
    for (int i = 0; i < P; i++)
        if (h->pairs[i].dir >= 0 && h->pairs[i].scale < ss_min) ss_min = h->pairs[i].scale;
    for (int i = 0; i < P; i++) {
        const clt_pair& p = h->pairs[i];
        a.cam1[i] = p.cam1;
        a.cam2[i] = p.cam2;
        a.strip_slot[i] = -1;
        if (p.dir >= 0) {
            const int s = a.num_strip++;
            const int ss = p.scale / ss_min, denom = ss * ((p.dir > 1) ? 1 : 2);
            int lg = 0;
            while ((1 << lg) < denom) lg++;
            if ((1 << lg) != denom) return fail(h, CLT_E_UNSUPPORTED, "clt_exec_corr2d: pair %d scale ratio %d is not a power of 2", i, ss);
            a.strip_slot[i] = s;
            a.strip_dir[s] = p.dir;
            a.strip_lg[s] = lg;
            a.strip_is_hor1[s] = (p.dir == 0 && p.scale == 1);
            a.strip_is_vert1[s] = (p.dir == 1 && p.scale == 1);
            a.n_hor1 += a.strip_is_hor1[s];
            a.n_vert1 += a.strip_is_vert1[s];
        }
    }
    if (g.num_colors == 3) {   // ImageDtt.java:1571-1581
        const double gw = 1.0 / (1.0 + h->cp.corr_red + h->cp.corr_blue);
        a.col_w[2] = (float)gw;
        a.col_w[0] = (float)(h->cp.corr_red * gw);
        a.col_w[1] = (float)(h->cp.corr_blue * gw);
    } else {
        a.col_w[0] = 1.f;
    }
    a.fat_zero = h->cp.fat_zero;
    a.corr_offset = h->cp.corr_offset;
    a.min_corr = h->cp.min_corr;
    a.twice_diagonal = h->cp.twice_diagonal;
    a.strip_height = hw;
    half_flat_top(2, 1.0, 1.0, true, false, 1.0, a.wndy);   // IDP:77-79 corr_wndy_*
    half_flat_top(9, 6.0, 5.0, true, false, 2.0, a.wndx);   // IDP:80-82 corr_wndx_*, wndx_scale 2.0 (ID:1758)
    for (int i = 0; i < 15; i++) {                          // createOrtoNotch (C2:2654-2675)
        double f = ((i < (8 - h->cp.enhortho_width)) || (i > (8 - 2 + h->cp.enhortho_width))) ? 1.0 : h->cp.enhortho_scale;
        if (i == 7) f = 0.0;
        a.notch[i] = (float)(f * std::sin(M_PI * (i + 1.0) / 16));
    }
    bool q = g.num_cams == 4 && g.num_colors == 3 && P == 6;
    if (q) {
        clt_pair dq[6];
        clt_default_quad_pairs(dq);
        for (int i = 0; i < 6; i++)
            q = q && h->pairs[i].cam1 == dq[i].cam1 && h->pairs[i].cam2 == dq[i].cam2 && h->pairs[i].dir == dq[i].dir &&
                h->pairs[i].scale == 1;
    }
    *quad = q;
    return CLT_OK;
}

// Correlation arguments and output buffers for the current pairs/parameters (CorrArgs upload).
static int prepare_corr(clt_handle* h) {
    CorrArgs a;
    bool quad = false;
    int rc;
    if ((rc = fill_corr_args(h, a, &quad))) return rc;
    const int P = (int)h->pairs.size();
    if ((rc = ensure_capacity(h, &h->d_corr, &h->corr_cap, (size_t)h->cap_tasks * P * CLT_CORR_LEN)) ||
        (rc = ensure_capacity(h, &h->d_combo, &h->combo_cap, (size_t)h->cap_tasks * 16 * 15)) ||
        (rc = dalloc(h, (void**)&h->d_disp, (size_t)h->cap_tasks * CLT_DISP_LAYERS * 4)))
        return rc;
    HIPCHK(h, upload_args(h, h->d_corr_args, a, h->corr_args_dev, h->corr_args_valid));
    if ((rc = build_lma_args(h, a.num_strip > 0))) return rc;
    h->corr_quad = quad;
    h->corr_num_strip = a.num_strip;
    h->corr_pairs_at_exec = P;
    h->corr_strip_h_at_exec = a.num_strip ? h->cp.strip_height : 0;
    return CLT_OK;
}

// When the last LMA may still read the correlation records, the next correlate2D writes the other
// pair (it waits only for the LMA before that one, which read the other pair).
static void swap_corr_records(clt_handle* h) {
    if (!h->lma_pending) return;
    join_lma2(h);
    std::swap(h->d_corr, h->d_corr2);
    std::swap(h->corr_cap, h->corr_cap2);
    std::swap(h->d_disp, h->d_disp2);
    std::swap(h->ev_lma, h->ev_lma2);
    h->lma2_pending = true;
    h->lma_pending = false;
}

// execConverCorrectTiles then execCorrelate2D over the same task list.  In the quad layout (4
// cameras, 3 colours, the 6 default pairs) one kernel does both, the CLT passed through LDS (it still
// goes to memory for the textures and imclt); otherwise the two launches of clt_exec_convert and
// clt_exec_corr2d.
int clt_exec_convert_corr2d(clt_handle* h, int32_t lpf_mask) {
    if (!h) return CLT_E_ARG;
    for (int c = 0; c < h->g.num_cams; c++)
        if (!h->images_set[c]) return fail(h, CLT_E_STATE, "clt_exec_convert_corr2d: image of camera %d not set", c);
    if (h->num_tasks <= 0) return fail(h, CLT_E_STATE, "clt_exec_convert_corr2d: no tasks");
    if (h->num_frames > 1) return fail(h, CLT_E_UNSUPPORTED, "clt_exec_convert_corr2d: a batch handle runs clt_exec_batch");
    int rc;
    swap_corr_records(h);
    if ((rc = prepare_corr(h))) return rc;
    const int n = h->num_tasks;
    if (h->corr_quad) {
        join_side(h);   // a running imclt_rbg still reads the CLT
        join_lma(h);    // the previous LMA still reads d_corr and d_disp
        tbegin(h, T_CONV_CORR);
        launch_convert_corr_range(h, 0, n, nullptr, lpf_mask);
        tend(h, T_CONV_CORR);
    } else {
        tbegin(h, T_CONVERT);
        launch_convert(h, n, nullptr, lpf_mask);
        tend(h, T_CONVERT);
        tbegin(h, T_CORR);
        launch_corr(h, n, nullptr);
        tend(h, T_CORR);
    }
    launch_corr_lma(h, n, true);
    HIPCHK(h, hipGetLastError());
    h->converted = true;
    h->tex_done = h->combined = h->rig_done = false;
    h->corr_done = true;
    return finish(h);
}

int clt_exec_corr2d(clt_handle* h) {
    if (!h) return CLT_E_ARG;
    if (h->num_frames > 1) return fail(h, CLT_E_UNSUPPORTED, "clt_exec_corr2d: a batch handle runs clt_exec_batch");
    if (!h->converted) return fail(h, CLT_E_STATE, "clt_exec_corr2d: no converted tiles");
    int rc;
    swap_corr_records(h);
    if ((rc = prepare_corr(h))) return rc;
    tbegin(h, T_CORR);
    launch_corr(h, h->num_tasks, nullptr);
    tend(h, T_CORR);
    launch_corr_lma(h, h->num_tasks, true);
    HIPCHK(h, hipGetLastError());
    h->corr_done = true;
    return finish(h);
}

// ------------------------------------------------------------------------ two-quad rig
void clt_default_rig_params(clt_rig_params* p) {   // BiQuadParameters BQP:32-83, IDP:81-112, ECP:2566
    p->use_poly = 1;
    p->use_xy_poly = 1;
    p->no_int_x0 = 1;
    p->notch_mode = 0;
    p->lt_rad = 0;
    p->corr_strip_notch = 11;
    p->corr_notch_hwidth = 2.0;
    p->corr_notch_blur = 2.0;
    p->cm_max_normalization = 0.55;
    p->max_corr_radius = 3.9;
}

int clt_set_rig_params(clt_handle* h, const clt_rig_params* p) {
    if (!h || !p) return CLT_E_ARG;
    if (p->lt_rad < 0 || p->lt_rad > 8 || p->corr_strip_notch < 1 || p->corr_strip_notch > 64 ||
        !(fabs(p->max_corr_radius) <= 7.0) || !(p->cm_max_normalization > 0))
        return fail(h, CLT_E_ARG, "clt_set_rig_params: lt_rad %d (0..8), notch rows %d (1..64), radius %g (|r| <= 7)",
                    p->lt_rad, p->corr_strip_notch, p->max_corr_radius);
    h->rp = *p;
    return CLT_OK;
}

// Buffers, CorrArgs of the quad pairs, the rig's LmaArgs and RigArgs, the target disparities.
static int prepare_rig(clt_handle* h, const float* target) {
    const clt_geom& g = h->g;
    if (g.num_cams != 8 || g.num_colors != 3)
        return fail(h, CLT_E_UNSUPPORTED, "clt_exec_corr2d_rig: needs 8 cameras (two quads) of Bayer frames");
    const size_t T = h->cap_tasks;
    int rc;
    if ((rc = dalloc(h, (void**)&h->d_rig_corr, 2 * T * 6 * CLT_CORR_LEN * 4)) ||
        (rc = dalloc(h, (void**)&h->d_rig_disp, 2 * T * CLT_DISP_LAYERS * 4)) ||
        (rc = dalloc(h, (void**)&h->d_rig_mix, 2 * T * 3 * 256 * 4)) ||
        (rc = dalloc(h, (void**)&h->d_rig_dirs, 2 * T * 8 * 4)) ||
        (rc = dalloc(h, (void**)&h->d_inter, T * CLT_CORR_LEN * 4)) ||
        (rc = dalloc(h, (void**)&h->d_inter_rec, T * RIG_REC * 4)) ||
        (rc = dalloc(h, (void**)&h->d_inter_res, T * 4 * 4)) ||
        (rc = dalloc(h, (void**)&h->d_bimap, T * CLT_BI_LAYERS * 4)) ||
        (rc = dalloc(h, (void**)&h->d_rig_extra, 2 * T * 4)) ||
        (rc = dalloc(h, (void**)&h->d_rig_args, 2 * sizeof(RigArgs))) ||
        (rc = dalloc(h, (void**)&h->d_lma_rig, sizeof(LmaArgs))) || (rc = ensure_lma_scratch(h, 0)) ||
        (rc = ensure_lma_scratch(h, 1)) ||
        (h->rp.lt_rad > 0 && (rc = dalloc(h, (void**)&h->d_inter_avg, T * CLT_CORR_LEN * 4))) ||
        (target && (rc = dalloc(h, (void**)&h->d_rig_target, T * 4))))
        return rc;
    const clt_rig_params& rp = h->rp;
    const int hw = h->cp.strip_height;
    // CorrArgs: the quad pair table on cameras 0..3 of each quad's CLT base
    std::vector<clt_pair> quad(6);
    clt_default_quad_pairs(quad.data());
    CorrArgs a;
    std::memset(&a, 0, sizeof(a));
    a.num_pairs = 6;
    for (int i = 0; i < 6; i++) {
        a.cam1[i] = quad[i].cam1;
        a.cam2[i] = quad[i].cam2;
        a.strip_slot[i] = i;
        a.strip_dir[i] = quad[i].dir;
        a.strip_lg[i] = quad[i].dir > 1 ? 0 : 1;
        a.strip_is_hor1[i] = quad[i].dir == 0;
        a.strip_is_vert1[i] = quad[i].dir == 1;
    }
    a.num_strip = 6;
    a.n_hor1 = a.n_vert1 = 2;
    const double gw = 1.0 / (1.0 + h->cp.corr_red + h->cp.corr_blue);   // ImageDtt.java:8793-8796
    a.col_w[2] = (float)gw;
    a.col_w[0] = (float)(h->cp.corr_red * gw);
    a.col_w[1] = (float)(h->cp.corr_blue * gw);
    a.fat_zero = h->cp.fat_zero;
    a.corr_offset = h->cp.corr_offset;
    a.min_corr = h->cp.min_corr;
    a.twice_diagonal = h->cp.twice_diagonal;
    a.strip_height = hw;
    half_flat_top(2, 1.0, 1.0, true, false, 1.0, a.wndy);
    half_flat_top(9, 6.0, 5.0, true, false, 2.0, a.wndx);
    for (int i = 0; i < 15; i++) {
        double f = ((i < (8 - h->cp.enhortho_width)) || (i > (8 - 2 + h->cp.enhortho_width))) ? 1.0 : h->cp.enhortho_scale;
        if (i == 7) f = 0.0;
        a.notch[i] = (float)(f * std::sin(M_PI * (i + 1.0) / 16));
    }
    // the tileCorrs LMA: mixing by rig.use_poly, the four directions on every fitted tile (dir_out),
    // no foreground correction; fit LMA_FIT_SINGLE is one correlation as a horizontal pair
    LmaArgs la;
    if ((rc = fill_lma_args(h, quad, h->lp, la))) return rc;
    la.mix = rp.use_poly;
    la.fo_correct = 0;
    la.fit[LMA_FIT_SINGLE].ngroups = 1;
    la.fit[LMA_FIT_SINGLE].g[0].npairs = 1;
    la.fit[LMA_FIT_SINGLE].g[0].diag = 0;
    la.fit[LMA_FIT_SINGLE].g[0].scale = 1.0;
    // RigArgs: tileInterCamCorrs as set, and its low-texture call with no_int_x0 (ID:9455-9466)
    RigArgs ra[2];
    for (int k = 0; k < 2; k++) {
        RigArgs& r = ra[k];
        std::memset(&r, 0, sizeof(r));
        const bool nix = k ? true : rp.no_int_x0 != 0;
        r.min_corr = (float)((nix ? 0.5 : 1.0) * h->cp.min_corr);
        r.cm_at_zero = nix || rp.notch_mode;
        // getMaxXCm (C2:936-1032) / getMaxXCmNotch (C2:925-934) windows of the Correlation2d
        // constructor (C2:139-159)
        const int wlen = rp.notch_mode ? rp.corr_strip_notch : 2;
        std::vector<float> wy(wlen);
        if (rp.notch_mode) half_flat_top(wlen, rp.corr_notch_hwidth, rp.corr_notch_blur, true, true, 2.0, wy.data());
        else half_flat_top(2, 1.0, 1.0, true, false, 1.0, wy.data());
        int rows = hw;
        double wy_scale = 1.0;
        if (rows > wlen) {
            rows = wlen;
        } else if (rows < wlen) {
            double swy = wy[0];
            for (int i = 1; i < rows; i++) swy += wy[i];
            wy_scale = 1.0 / swy;
        }
        r.cm_rows = rows;
        for (int i = 0; i < rows; i++) r.cm_wy[i] = (float)(((i == 0) ? wy_scale : 2.0 * wy_scale) * wy[i]);
        half_flat_top(9, 6.0, 5.0, true, false, 2.0, r.wndx);
        r.cm_max_norm = (float)rp.cm_max_normalization;
        r.iradius = (int)std::fabs(rp.max_corr_radius);
        r.ir2 = (int)(rp.max_corr_radius * rp.max_corr_radius);
        r.square = rp.max_corr_radius < 0;
    }
    join_lma(h);
    HIPCHK(h, upload_args(h, h->d_corr_args, a, h->corr_args_dev, h->corr_args_valid));
    HIPCHK(h, hipMemcpyAsync(h->d_lma_rig, &la, sizeof(la), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_rig_args, ra, sizeof(ra), hipMemcpyHostToDevice, h->stream));
    h->rig_target = target != nullptr;
    if (target) HIPCHK(h, hipMemcpyAsync(h->d_rig_target, target, (size_t)h->num_tasks * 4, hipMemcpyHostToDevice, h->stream));
    h->graph_epoch++;   // the correlation arguments now hold the quad table
    return CLT_OK;
}

int clt_exec_corr2d_rig(clt_handle* h, const float* target_disparity) {
    if (!h) return CLT_E_ARG;
    if (h->num_frames > 1) return fail(h, CLT_E_UNSUPPORTED, "clt_exec_corr2d_rig: a batch handle runs clt_exec_batch");
    if (!h->converted) return fail(h, CLT_E_STATE, "clt_exec_corr2d_rig: no converted tiles");
    int rc;
    if ((rc = prepare_rig(h, target_disparity))) return rc;
    const int n = h->num_tasks;
    const size_t T = h->cap_tasks, ntiles = (size_t)h->tiles_x * h->tiles_y;
    const bool fz = h->cp.fat_zero != 0.f;
    const int nc = h->lp.lma_adjust_wy ? 4 : 3;
    tbegin(h, T_RIG);
    const dim3 g4((n + 3) / 4), g_lma((n + LMA_SLOTS - 1) / LMA_SLOTS);
    for (int q = 0; q < 2; q++) {   // each quad's correlate2D, with its mean CLT
        const int tpb = h->tune_corr_tpb;
        auto kern = fz ? k_corr2d_quad<true> : k_corr2d_quad<false>;
        hipLaunchKernelGGL(kern, dim3((n + tpb - 1) / tpb), dim3(64 * tpb), (size_t)tpb * CQ_WAVE_FLOATS * 4, h->stream,
                           h->d_clt + (size_t)q * 4 * ntiles * 3 * 256, h->d_tasks, h->task_words, n, h->d_lpf_corr,
                           h->d_corr_args, h->d_rig_corr + q * T * 6 * CLT_CORR_LEN, nullptr,
                           h->d_rig_disp + q * T * CLT_DISP_LAYERS, nullptr, 0LL, h->gd, h->d_rig_mix + q * T * 3 * 256);
        if (q == 0) {   // the main quad's LMA on the second stream, beside the aux quad's correlate2D
            HIPCHK(h, hipEventRecord(h->ev_fork, h->stream));
            HIPCHK(h, hipStreamWaitEvent(h->stream_lma, h->ev_fork, 0));
            launch_lma_chain(h, h->stream_lma, 0, nc, h->d_rig_corr, h->d_rig_disp, n, h->d_lma_rig, h->d_rig_dirs);
            HIPCHK(h, hipEventRecord(h->ev_lma, h->stream_lma));
        }
    }
    auto inter = fz ? k_corr_inter<true, true> : k_corr_inter<false, true>;
    hipLaunchKernelGGL(inter, g4, dim3(256), 0, h->stream, h->d_rig_mix, h->d_rig_mix + T * 3 * 256, nullptr, h->d_tasks,
                       h->task_words, n, h->d_lpf_corr, h->d_corr_args, h->d_rig_args, h->d_inter, h->d_inter_rec);
    // the aux quad's LMA on the side stream, beside the main quad's and the inter-camera kernels
    // (they read only the two quads' mean CLT and their own buffers)
    join_side(h);
    HIPCHK(h, hipEventRecord(h->ev_side_fork, h->stream));
    HIPCHK(h, hipStreamWaitEvent(h->stream_side, h->ev_side_fork, 0));
    launch_lma_chain(h, h->stream_side, 1, nc, h->d_rig_corr + T * 6 * CLT_CORR_LEN, h->d_rig_disp + T * CLT_DISP_LAYERS, n,
                     h->d_lma_rig, h->d_rig_dirs + T * 8);
    HIPCHK(h, hipEventRecord(h->ev_side, h->stream_side));
    h->side_pending = true;
    auto lma_x = (nc == 4) ? k_corr_lma_inter<4> : k_corr_lma_inter<3>;
    const clt_rig_params& rp = h->rp;
    hipLaunchKernelGGL(lma_x, g_lma, dim3(64 * LMA_WAVES), 0, h->stream, h->d_inter, h->d_inter_rec, h->d_inter_res, n,
                       h->d_lma_rig, rp.use_poly, rp.use_xy_poly, rp.notch_mode);
    if (rp.lt_rad > 0) {   // low-texture pass (ID:9404-9480)
        hipLaunchKernelGGL(k_inter_lt_avg, g4, dim3(256), 0, h->stream, h->d_inter, h->d_tasks, h->task_words, n,
                           h->d_tile_map, rp.lt_rad, rp.notch_mode ? 0 : rp.lt_rad, h->gd, h->d_inter_avg);
        auto inter_lt = k_corr_inter<false, false>;   // no mix input: the fat zero is not read
        hipLaunchKernelGGL(inter_lt, g4, dim3(256), 0, h->stream, nullptr, nullptr, h->d_inter_avg, h->d_tasks,
                           h->task_words, n, h->d_lpf_corr, h->d_corr_args, h->d_rig_args + 1, nullptr, h->d_inter_rec);
        hipLaunchKernelGGL(lma_x, g_lma, dim3(64 * LMA_WAVES), 0, h->stream, h->d_inter_avg, h->d_inter_rec,
                           h->d_inter_res, n, h->d_lma_rig, rp.use_poly, rp.use_xy_poly, rp.notch_mode);
    }
    HIPCHK(h, hipStreamWaitEvent(h->stream, h->ev_lma, 0));   // the main quad's LMA
    join_side(h);                                              // the aux quad's LMA
    hipLaunchKernelGGL(k_rig_bimap, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->d_tasks, h->task_words, n,
                       h->d_rig_disp, h->d_rig_dirs, h->d_rig_disp + T * CLT_DISP_LAYERS, h->d_rig_dirs + T * 8,
                       h->d_inter_res, h->rig_target ? h->d_rig_target : nullptr, h->d_bimap, h->d_rig_extra,
                       h->d_rig_extra + T);
    tend(h, T_RIG);
    HIPCHK(h, hipGetLastError());
    h->rig_done = true;
    h->rig_tex_done = false;
    return finish(h);
}

int clt_get_corr2d_rig(clt_handle* h, float* bimap, float* inter, float* corr_main, float* corr_aux, float* disp_main,
                       float* disp_aux) {
    if (!h) return CLT_E_ARG;
    if (!h->rig_done) return fail(h, CLT_E_STATE, "clt_get_corr2d_rig: rig correlation not executed");
    HIPCHK(h, sync_streams(h));
    const size_t nt = h->num_tasks, T = h->cap_tasks;
    if (bimap) HIPCHK(h, hipMemcpy(bimap, h->d_bimap, nt * CLT_BI_LAYERS * 4, hipMemcpyDeviceToHost));
    if (inter) HIPCHK(h, hipMemcpy(inter, h->d_inter, nt * CLT_CORR_LEN * 4, hipMemcpyDeviceToHost));
    float* dst[4] = {corr_main, corr_aux, disp_main, disp_aux};
    for (int q = 0; q < 2; q++) {
        if (dst[q])
            HIPCHK(h, hipMemcpy(dst[q], h->d_rig_corr + q * T * 6 * CLT_CORR_LEN, nt * 6 * CLT_CORR_LEN * 4,
                                hipMemcpyDeviceToHost));
        if (dst[2 + q])
            HIPCHK(h, hipMemcpy(dst[2 + q], h->d_rig_disp + q * T * CLT_DISP_LAYERS, nt * CLT_DISP_LAYERS * 4,
                                hipMemcpyDeviceToHost));
    }
    return CLT_OK;
}

// Texture arguments and output buffers for the current parameters (TexArgs upload).
// TexArgs of ncams ports with the given port offsets and the handle's texture parameters.
static void fill_tex_args(const clt_handle* h, int ncams, const float (*port_offsets)[2], TexArgs& a) {
    std::memset(&a, 0, sizeof(a));
    const double gw = 1.0 / (1.0 + h->cp.corr_red + h->cp.corr_blue);
    a.col_w[2] = (float)gw;
    a.col_w[0] = (float)(h->cp.corr_red * gw);
    a.col_w[1] = (float)(h->cp.corr_blue * gw);
    a.min_shot = h->tp.min_shot;
    a.scale_shot = h->tp.scale_shot;
    const double ksigma = 1.0 / (2.0 * h->tp.diff_sigma * h->tp.diff_sigma);
    a.ksigma = (float)ksigma;
    a.ksigma_l2 = (float)(ksigma * M_LOG2E);
    for (int i = 0, np = 0; i < ncams; i++)
        for (int j = i + 1; j < ncams; j++, np++) {   // ImageDtt.java:2881-2891 pair weights
            const double dx = (double)port_offsets[j][0] - port_offsets[i][0];
            const double dy = (double)port_offsets[j][1] - port_offsets[i][1];
            a.pair_k_l2[np] = (float)(ksigma / (dx * dx + dy * dy) * M_LOG2E);
        }
    a.corr_magic_scale = h->tp.corr_magic_scale;
    a.dust_remove = h->tp.dust_remove;
    a.planes = h->tp.keep_weights ? 4 + ncams + 4 : 4;
    for (int i = 0; i < ncams; i++) {
        a.port_offsets[i][0] = port_offsets[i][0];
        a.port_offsets[i][1] = port_offsets[i][1];
    }
}

static int prepare_tex(clt_handle* h) {
    const clt_geom& g = h->g;
    if (g.num_colors != 3 || g.num_cams < 2 || g.num_cams > 8)
        return fail(h, CLT_E_UNSUPPORTED, "clt_exec_textures: needs a Bayer frame of 2..8 cameras");
    const int planes = h->tp.keep_weights ? 4 + g.num_cams + 4 : 4;
    int rc;
    if ((rc = ensure_capacity(h, &h->d_tex, &h->tex_cap, (size_t)h->cap_tasks * planes * 256)) ||
        (rc = dalloc(h, (void**)&h->d_tex_stats, (size_t)h->cap_tasks * 4 * CLT_MAX_CAMS * 4)))
        return rc;
    TexArgs a;
    fill_tex_args(h, g.num_cams, h->tp.port_offsets, a);
    HIPCHK(h, upload_args(h, h->d_tex_args, a, h->tex_args_dev, h->tex_args_valid));
    h->tex_planes_at_exec = planes;
    return CLT_OK;
}

int clt_exec_textures(clt_handle* h, const float* extra_disparity) {
    if (!h) return CLT_E_ARG;
    if (h->num_frames > 1) return fail(h, CLT_E_UNSUPPORTED, "clt_exec_textures: a batch handle runs clt_exec_batch");
    if (!h->converted) return fail(h, CLT_E_STATE, "clt_exec_textures: no converted tiles");
    int rc;
    if ((rc = prepare_tex(h))) return rc;
    const float* extra = nullptr;
    if (extra_disparity) {
        HIPCHK(h, hipMemcpyAsync(h->d_extra, extra_disparity, (size_t)h->num_tasks * 4, hipMemcpyHostToDevice, h->stream));
        extra = h->d_extra;
    }
    tbegin(h, T_TEX);
    launch_tex(h, h->num_tasks, nullptr, extra);
    tend(h, T_TEX);
    HIPCHK(h, hipGetLastError());
    h->tex_done = true;
    h->combined = false;
    return finish(h);
}

int clt_exec_textures_imclt(clt_handle* h, const float* extra_disparity, int32_t apply_lpf) {
    if (!h) return CLT_E_ARG;
    if (h->num_frames > 1) return fail(h, CLT_E_UNSUPPORTED, "clt_exec_textures_imclt: a batch handle runs clt_exec_batch");
    if (!h->converted) return fail(h, CLT_E_STATE, "clt_exec_textures_imclt: no converted tiles");
    const clt_geom& g = h->g;
    const bool fuse = !extra_disparity && apply_lpf && h->tex_fusable && g.num_colors == 3 && g.num_cams >= 2 &&
                      g.num_cams <= 5;
    if (!fuse) {   // the two passes
        int rc = clt_exec_textures(h, extra_disparity);
        return rc ? rc : clt_exec_imclt(h, apply_lpf);
    }
    int rc;
    if ((rc = prepare_tex(h))) return rc;
    if ((rc = dalloc(h, (void**)&h->d_rbg, (size_t)g.num_cams * g.num_colors * (g.width + 8) * (g.height + 8) * 4)))
        return rc;
    join_side(h);   // an imclt_rbg still writing the planes
    tbegin(h, T_TEX_IMCLT);
    launch_tex_imclt(h);
    tend(h, T_TEX_IMCLT);
    HIPCHK(h, hipGetLastError());
    h->tex_done = true;
    h->combined = false;
    return finish(h);
}

// generateTextureTiles of both quads of the rig (ID:9360-9398): each quad's 4 ports with the quad
// port offsets (ID:8832-8836) and its tileCorrs disparity as the extra disparity.
int clt_exec_textures_rig(clt_handle* h) {
    if (!h) return CLT_E_ARG;
    if (!h->rig_done) return fail(h, CLT_E_STATE, "clt_exec_textures_rig: rig correlation not executed");
    const size_t T = h->cap_tasks, ntiles = (size_t)h->tiles_x * h->tiles_y;
    static const float quad_offsets[4][2] = {{-0.5f, -0.5f}, {0.5f, -0.5f}, {-0.5f, 0.5f}, {0.5f, 0.5f}};
    TexArgs a;
    fill_tex_args(h, 4, quad_offsets, a);
    const int planes = a.planes;
    int rc;
    if ((rc = dalloc(h, (void**)&h->d_rig_tex_args, sizeof(TexArgs))) ||
        (rc = dalloc(h, (void**)&h->d_rig_tex_stats, 2 * T * 16 * 4)))
        return rc;
    if (planes > h->rig_tex_planes) {   // the buffer grows when keep_weights is switched on
        if (h->d_rig_tex) {
            HIPCHK(h, sync_streams(h));   // the previous rig textures may still be written
            HIPCHK(h, hipFree(h->d_rig_tex));
        }
        h->d_rig_tex = nullptr;
        if ((rc = dalloc(h, (void**)&h->d_rig_tex, 2 * T * planes * 256 * 4))) return rc;
        h->rig_tex_planes = planes;
    }
    HIPCHK(h, hipMemcpyAsync(h->d_rig_tex_args, &a, sizeof(a), hipMemcpyHostToDevice, h->stream));
    const int n = h->num_tasks;
    tbegin(h, T_TEX);
    for (int q = 0; q < 2; q++)
        hipLaunchKernelGGL(k_textures<4>, dim3(n), dim3(256), 0, h->stream, h->d_clt + (size_t)q * 4 * ntiles * 3 * 256,
                           h->d_tasks, h->task_words, n, h->d_lpf_corr, h->d_rig_extra + q * T, h->d_rig_tex_args,
                           h->d_rig_tex + q * T * planes * 256, h->d_rig_tex_stats + q * T * 16, nullptr, 0LL, h->gd);
    tend(h, T_TEX);
    HIPCHK(h, hipGetLastError());
    h->rig_tex_done = true;
    h->tex_planes_rig = planes;
    return finish(h);
}

int clt_get_textures_rig(clt_handle* h, float* tiles_main, float* tiles_aux, float* stats_main, float* stats_aux) {
    if (!h) return CLT_E_ARG;
    if (!h->rig_tex_done) return fail(h, CLT_E_STATE, "clt_get_textures_rig: rig textures not executed");
    HIPCHK(h, sync_streams(h));
    const size_t nt = h->num_tasks, T = h->cap_tasks, planes = h->tex_planes_rig;
    float* tl[2] = {tiles_main, tiles_aux};
    float* st[2] = {stats_main, stats_aux};
    for (int q = 0; q < 2; q++) {
        if (tl[q])
            HIPCHK(h, hipMemcpy(tl[q], h->d_rig_tex + q * T * planes * 256, nt * planes * 256 * 4, hipMemcpyDeviceToHost));
        if (st[q]) HIPCHK(h, hipMemcpy(st[q], h->d_rig_tex_stats + q * T * 16, nt * 16 * 4, hipMemcpyDeviceToHost));
    }
    return CLT_OK;
}

int clt_exec_combine_textures(clt_handle* h) { return clt_exec_combine_rgba_tiles(h, 1, 0); }

int clt_exec_combine_rgba_tiles(clt_handle* h, int32_t overlap, int32_t sharp_alpha) {
    if (!h) return CLT_E_ARG;
    if (h->num_frames > 1) return fail(h, CLT_E_UNSUPPORTED, "clt_exec_combine_rgba_tiles: not available for a batch");
    if (!h->tex_done) return fail(h, CLT_E_STATE, "clt_exec_combine_rgba_tiles: no textures");
    const clt_geom& g = h->g;
    const int planes = h->tex_planes_at_exec;
    const int k = overlap ? 1 : 2;
    int rc;
    if ((rc = ensure_capacity(h, &h->d_tex_img, &h->tex_img_cap, (size_t)planes * k * g.width * k * g.height))) return rc;
    dim3 grid((k * g.width + 63) / 64, (k * g.height + 3) / 4);
    tbegin(h, T_COMBINE);
    hipLaunchKernelGGL(k_combine_textures, grid, dim3(256), 0, h->stream, h->d_tex, h->d_tile_map, planes, overlap ? 1 : 0,
                       sharp_alpha ? 1 : 0, h->d_tex_img, h->gd);
    tend(h, T_COMBINE);
    HIPCHK(h, hipGetLastError());
    h->combined = true;
    h->combine_scale = k;
    return finish(h);
}

// ------------------------------------------------------------------------ outputs
int clt_get_rbg(clt_handle* h, int32_t cam, float* out) {
    if (!h || !out) return CLT_E_ARG;
    if (cam < 0 || cam >= h->g.num_cams) return fail(h, CLT_E_ARG, "clt_get_rbg: camera %d", cam);
    if (!h->d_rbg) return fail(h, CLT_E_STATE, "clt_get_rbg: imclt not executed");
    const size_t n = (size_t)h->g.num_colors * (h->g.width + 8) * (h->g.height + 8);
    HIPCHK(h, sync_streams(h));
    HIPCHK(h, hipMemcpy(out, h->d_rbg + cam * n, n * 4, hipMemcpyDeviceToHost));
    return CLT_OK;
}

int clt_get_clt(clt_handle* h, int32_t cam, float* out) {
    if (!h || !out) return CLT_E_ARG;
    if (cam < 0 || cam >= h->g.num_cams) return fail(h, CLT_E_ARG, "clt_get_clt: camera %d", cam);
    const size_t n = (size_t)h->num_tiles * h->g.num_colors * 256;
    std::vector<float> tmp(n);
    HIPCHK(h, sync_streams(h));
    HIPCHK(h, hipMemcpy(tmp.data(), h->d_clt + cam * n, n * 4, hipMemcpyDeviceToHost));
    for (size_t b = 0; b < n / 64; b++)   // device [h][v] -> [v][h] (TileProcessor.cuh:1065-1095 layout)
        for (int hh = 0; hh < 8; hh++)
            for (int v = 0; v < 8; v++) out[b * 64 + v * 8 + hh] = tmp[b * 64 + hh * 8 + v];
    return CLT_OK;
}

int clt_get_corr_offsets(clt_handle* h, double radius, float* offsets) {
    if (!h || !offsets) return CLT_E_ARG;
    if (!(radius >= 1.0 && radius <= 7.0)) return fail(h, CLT_E_ARG, "clt_get_corr_offsets: radius %g outside [1, 7]", radius);
    if (!h->corr_done) return fail(h, CLT_E_STATE, "clt_get_corr_offsets: corr2d not executed");
    if (!h->corr_quad || h->corr_pairs_at_exec != 6)
        return fail(h, CLT_E_UNSUPPORTED, "clt_get_corr_offsets: needs the 6 reference quad pairs");
    const int n = h->num_tasks;
    if (n == 0) return CLT_OK;
    int rc;
    if ((rc = dalloc(h, (void**)&h->d_offs, (size_t)h->cap_tasks * 18 * 4))) return rc;
    hipLaunchKernelGGL(k_corr_offsets, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->d_corr, h->d_tasks,
                       h->task_words, n, (float)radius, 0.f, nullptr, h->d_offs);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, sync_streams(h));
    HIPCHK(h, hipMemcpy(offsets, h->d_offs, (size_t)n * 18 * 4, hipMemcpyDeviceToHost));
    return CLT_OK;
}

int clt_get_corr2d(clt_handle* h, float* corr, float* combo, float* disparity) {
    if (!h) return CLT_E_ARG;
    if (!h->corr_done) return fail(h, CLT_E_STATE, "clt_get_corr2d: corr2d not executed");
    HIPCHK(h, sync_streams(h));
    const size_t nt = h->num_tasks;
    if (corr) HIPCHK(h, hipMemcpy(corr, h->d_corr, nt * h->corr_pairs_at_exec * CLT_CORR_LEN * 4, hipMemcpyDeviceToHost));
    if (combo && h->corr_strip_h_at_exec)
        HIPCHK(h, hipMemcpy(combo, h->d_combo, nt * h->corr_strip_h_at_exec * 15 * 4, hipMemcpyDeviceToHost));
    if (disparity && h->corr_strip_h_at_exec)
        HIPCHK(h, hipMemcpy(disparity, h->d_disp, nt * CLT_DISP_LAYERS * 4, hipMemcpyDeviceToHost));
    return CLT_OK;
}

int clt_get_textures(clt_handle* h, float* tiles, float* stats) {
    if (!h) return CLT_E_ARG;
    if (!h->tex_done) return fail(h, CLT_E_STATE, "clt_get_textures: textures not executed");
    HIPCHK(h, sync_streams(h));
    const size_t nt = h->num_tasks;
    if (tiles) HIPCHK(h, hipMemcpy(tiles, h->d_tex, nt * h->tex_planes_at_exec * 256 * 4, hipMemcpyDeviceToHost));
    if (stats) HIPCHK(h, hipMemcpy(stats, h->d_tex_stats, nt * 4 * h->g.num_cams * 4, hipMemcpyDeviceToHost));
    return CLT_OK;
}

int clt_get_combined_textures(clt_handle* h, float* out) {
    if (!h || !out) return CLT_E_ARG;
    if (!h->combined) return fail(h, CLT_E_STATE, "clt_get_combined_textures: combine not executed");
    HIPCHK(h, sync_streams(h));
    const size_t k = h->combine_scale;
    HIPCHK(h, hipMemcpy(out, h->d_tex_img, (size_t)h->tex_planes_at_exec * k * h->g.width * k * h->g.height * 4,
                        hipMemcpyDeviceToHost));
    return CLT_OK;
}

// ------------------------------------------------------------------------ correlation-peak loss adjoint
static int check_adjoint_config(clt_handle* h, const char* who) {
    const clt_geom& g = h->g;
    if (h->num_frames > 1 || g.num_cams != 4 || g.num_colors != 3)
        return fail(h, CLT_E_UNSUPPORTED, "%s: needs a single-frame quad Bayer handle", who);
    if (h->cp.fat_zero != 0.f) return fail(h, CLT_E_UNSUPPORTED, "%s: the adjoint is written for fat_zero 0", who);
    for (int c = 0; c < g.num_cams; c++)
        if (!h->images_set[c]) return fail(h, CLT_E_STATE, "%s: image of camera %d not set", who, c);
    if (h->num_tasks <= 0) return fail(h, CLT_E_STATE, "%s: no tasks", who);
    return CLT_OK;
}

static void launch_corr_adjoint(clt_handle* h, float* out) {
    hipLaunchKernelGGL(k_corr_adjoint_quad, dim3(h->num_tasks), dim3(192), 0, h->stream, h->d_images,
                       h->use_kernels ? h->d_kernels : nullptr, h->use_kernels ? h->d_koffs : nullptr, h->d_tasks,
                       h->task_words, h->num_tasks, h->d_lpf_corr, h->d_corr_args, 1.0 / h->g.kernel_step, out, h->gd);
}

int clt_exec_corr_adjoint(clt_handle* h) {
    if (!h) return CLT_E_ARG;
    int rc;
    if ((rc = check_adjoint_config(h, "clt_exec_corr_adjoint"))) return rc;
    if ((rc = prepare_corr(h))) return rc;   // colour weights
    if (!h->corr_quad) return fail(h, CLT_E_UNSUPPORTED, "clt_exec_corr_adjoint: needs the 6 reference quad pairs");
    if ((rc = dalloc(h, (void**)&h->d_adj, (size_t)h->cap_tasks * CORR_ADJ_STRIDE * 4))) return rc;
    tbegin(h, T_ADJ);
    launch_corr_adjoint(h, h->d_adj);
    tend(h, T_ADJ);
    HIPCHK(h, hipGetLastError());
    h->adj_done = true;
    return finish(h);
}

int clt_get_corr_adjoint(clt_handle* h, float* out) {
    if (!h || !out) return CLT_E_ARG;
    if (!h->adj_done) return fail(h, CLT_E_STATE, "clt_get_corr_adjoint: not executed");
    HIPCHK(h, sync_streams(h));
    HIPCHK(h, hipMemcpy(out, h->d_adj, (size_t)h->num_tasks * CORR_ADJ_STRIDE * 4, hipMemcpyDeviceToHost));
    return CLT_OK;
}

// ------------------------------------------------------------------------ geometry adjoint and LMA
void clt_default_lma_params(clt_lma_params* p) {
    if (!p) return;
    p->lambda = 0.1;               // IDP lma_lambda_initial
    p->lambda_scale_good = 0.5;    // IDP lma_lambda_scale_good
    p->lambda_scale_bad = 8.0;     // IDP lma_lambda_scale_bad
    p->lambda_max = 100.0;         // IDP lma_lambda_max
    p->rms_diff = 1e-6;
    p->num_iter = 10;
    p->adjust_disparity = 1;
    p->coarse_iter = 0;
    p->reserved = 0;
    p->coarse_radius = 3.9;      // ECP:2566 max_corr_radius
    p->coarse_strength_floor = 0.3;
}

static int geo_prepare(clt_handle* h, const char* who) {
    int rc;
    if ((rc = check_adjoint_config(h, who))) return rc;
    if (!h->port_geom_set || !h->tasks_from_geom)
        return fail(h, CLT_E_STATE, "%s: tasks must come from clt_set_tasks_geometry", who);
    if ((rc = prepare_corr(h))) return rc;
    if (!h->corr_quad) return fail(h, CLT_E_UNSUPPORTED, "%s: needs the 6 reference quad pairs", who);
    const int blocks = (h->max_tasks + GEO_ADJ_TASKS - 1) / GEO_ADJ_TASKS;
    if ((rc = dalloc(h, (void**)&h->d_adj, (size_t)h->cap_tasks * CORR_ADJ_STRIDE * 4)) ||
        (rc = dalloc(h, (void**)&h->d_tstate, (size_t)h->max_tasks * 15 * 8)) ||
        (rc = dalloc(h, (void**)&h->d_gpart, (size_t)blocks * GEO_ADJ_SUMS * 8)) ||
        (rc = dalloc(h, (void**)&h->d_gsum, (size_t)(GEO_ADJ_SUMS + 13) * 8)) ||
        (rc = dalloc(h, (void**)&h->d_gdisp2, (size_t)h->max_tasks * 4)))
        return rc;
    return CLT_OK;
}

// Port centres of the tasks at disparities `disp` with the current device geometry, then the loss
// adjoint; with `sums`, the geometry adjoint with damping lambda and its reduction to the host.
static int geo_eval(clt_handle* h, const float* disp, double lambda, double* sums) {
    const int n = h->num_tasks;
    hipLaunchKernelGGL(k_port_centres, dim3((4 * n + 255) / 256), dim3(256), 0, h->stream, h->d_port_geom, h->d_rtab,
                       h->d_gwords, h->d_gtxy, disp, n, h->task_words, h->d_tasks, nullptr, 0, 1.0);
    launch_corr_adjoint(h, h->d_adj);
    if (!sums) return CLT_OK;
    const int blocks = (n + GEO_ADJ_TASKS - 1) / GEO_ADJ_TASKS;
    hipLaunchKernelGGL(k_geometry_adjoint, dim3(blocks), dim3(128), 0, h->stream, h->d_port_geom, h->d_rtab, h->d_gwords,
                       h->d_gtxy, disp, h->d_adj, n, lambda, h->d_gpart, h->d_tstate);
    hipLaunchKernelGGL(k_sum_rows, dim3(GEO_ADJ_SUMS), dim3(256), 0, h->stream, h->d_gpart, blocks, GEO_ADJ_SUMS, h->d_gsum);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(sums, h->d_gsum, GEO_ADJ_SUMS * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, sync_streams(h));
    return CLT_OK;
}

static int upload_geometry(clt_handle* h, const clt_port_geometry& p) {
    PortGeomDev pg;
    port_geom_dev(&p, pg);
    HIPCHK(h, hipMemcpyAsync(h->d_port_geom, &pg, sizeof(pg), hipMemcpyHostToDevice, h->stream));
    return CLT_OK;
}

int clt_exec_geometry_adjoint(clt_handle* h, double* grad_extrinsic, float* grad_disparity, double* loss) {
    if (!h) return CLT_E_ARG;
    int rc;
    if ((rc = geo_prepare(h, "clt_exec_geometry_adjoint"))) return rc;
    double sums[GEO_ADJ_SUMS];
    tbegin(h, T_GEOADJ);
    if ((rc = geo_eval(h, h->d_gdisp, 0.0, sums))) return rc;
    tend(h, T_GEOADJ);
    h->adj_done = true;
    if (grad_extrinsic)
        for (int i = 0; i < 13; i++) grad_extrinsic[i] = sums[105 + i];
    if (loss) *loss = sums[104];
    if (grad_disparity && h->num_tasks > 0) {
        std::vector<double> ts((size_t)h->num_tasks * 15);
        HIPCHK(h, hipMemcpy(ts.data(), h->d_tstate, ts.size() * 8, hipMemcpyDeviceToHost));
        for (int t = 0; t < h->num_tasks; t++) grad_disparity[t] = (float)ts[(size_t)t * 15];
    }
    return CLT_OK;
}

// clt_exec_geometry_adjoint on the static scratch: no allocation on any call.  The handle's
// inputs (images, kernels, port geometry, task list of clt_set_tasks_geometry) and its CorrArgs
// buffer were allocated by earlier calls; the per-task and per-block buffers live in the scratch
// array, which every call on this GPU shares (each call rewrites all of it that it reads, so
// nothing has to start at zero; calls that overlap on one GPU race).
int clt_exec_geometry_adjoint_noalloc(clt_handle* h, double* grad_extrinsic, float* grad_disparity, double* loss) {
    if (!h) return CLT_E_ARG;
    const char* who = "clt_exec_geometry_adjoint_noalloc";
    int rc;
    if ((rc = check_adjoint_config(h, who))) return rc;
    if (!h->port_geom_set || !h->tasks_from_geom)
        return fail(h, CLT_E_STATE, "%s: tasks must come from clt_set_tasks_geometry", who);
    const int n = h->num_tasks;
    if (n > CLT_NOALLOC_MAX_TASKS)
        return fail(h, CLT_E_ARG, "%s: %d tasks, the static scratch holds at most %d", who, n, CLT_NOALLOC_MAX_TASKS);
    if (!h->scratch) return fail(h, CLT_E_STATE, "%s: no scratch address on device %d", who, h->device);
    CorrArgs a;
    bool quad = false;
    if ((rc = fill_corr_args(h, a, &quad))) return rc;
    if (!quad) return fail(h, CLT_E_UNSUPPORTED, "%s: needs the 6 reference quad pairs", who);
    unsigned char* base = (unsigned char*)h->scratch;
    float* adj = (float*)(base + NA_OFF_ADJ);
    double* tstate = (double*)(base + NA_OFF_TSTATE);
    double* part = (double*)(base + NA_OFF_PART);
    double* gsum = (double*)(base + NA_OFF_SUM);
    float* gdisp = (float*)(base + NA_OFF_GD);
    const int blocks = (n + GEO_ADJ_TASKS - 1) / GEO_ADJ_TASKS;
    join_lma(h);   // kernels of earlier calls on the other streams finish first
    join_side(h);
    HIPCHK(h, upload_args(h, h->d_corr_args, a, h->corr_args_dev, h->corr_args_valid));
    tbegin(h, T_GEOADJ);
    hipLaunchKernelGGL(k_port_centres, dim3((4 * n + 255) / 256), dim3(256), 0, h->stream, h->d_port_geom, h->d_rtab,
                       h->d_gwords, h->d_gtxy, h->d_gdisp, n, h->task_words, h->d_tasks, nullptr, 0, 1.0);
    launch_corr_adjoint(h, adj);
    hipLaunchKernelGGL(k_geometry_adjoint, dim3(blocks), dim3(128), 0, h->stream, h->d_port_geom, h->d_rtab, h->d_gwords,
                       h->d_gtxy, h->d_gdisp, adj, n, 0.0, part, tstate);
    hipLaunchKernelGGL(k_sum_rows, dim3(GEO_ADJ_SUMS), dim3(256), 0, h->stream, part, blocks, GEO_ADJ_SUMS, gsum);
    if (grad_disparity)
        hipLaunchKernelGGL(k_take_col, dim3((n + 255) / 256), dim3(256), 0, h->stream, tstate, 15, n, gdisp);
    tend(h, T_GEOADJ);
    HIPCHK(h, hipGetLastError());
    double sums[GEO_ADJ_SUMS];
    HIPCHK(h, hipMemcpyAsync(sums, gsum, sizeof(sums), hipMemcpyDeviceToHost, h->stream));
    if (grad_disparity && n > 0)
        HIPCHK(h, hipMemcpyAsync(grad_disparity, gdisp, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (grad_extrinsic)
        for (int i = 0; i < 13; i++) grad_extrinsic[i] = sums[105 + i];
    if (loss) *loss = sums[104];
    return CLT_OK;
}

// Cholesky solve of the 13 x 13 system (A + lambda diag(A)) x = -b; false if not positive definite.
static bool solve13(const double* sums, double lambda, double* x) {
    double a[13][13], l[13][13] = {};
    for (int i = 0, e = 0; i < 13; i++)
        for (int j = i; j < 13; j++, e++) a[i][j] = a[j][i] = sums[e];
    double dmax = 0.0;
    for (int i = 0; i < 13; i++) dmax = std::max(dmax, a[i][i]);
    for (int i = 0; i < 13; i++) a[i][i] += lambda * a[i][i] + 1e-12 * dmax + 1e-300;   // damping, gauge freedoms
    for (int i = 0; i < 13; i++) {
        for (int j = 0; j <= i; j++) {
            double s = a[i][j];
            for (int k = 0; k < j; k++) s -= l[i][k] * l[j][k];
            if (i == j) {
                if (!(s > 0.0)) return false;
                l[i][i] = std::sqrt(s);
            } else {
                l[i][j] = s / l[j][j];
            }
        }
    }
    double y[13];
    for (int i = 0; i < 13; i++) {
        double s = -sums[91 + i];
        for (int k = 0; k < i; k++) s -= l[i][k] * y[k];
        y[i] = s / l[i][i];
    }
    for (int i = 12; i >= 0; i--) {
        double s = y[i];
        for (int k = i + 1; k < 13; k++) s -= l[k][i] * x[k];
        x[i] = s / l[i][i];
    }
    return true;
}

int clt_lma_extrinsics(clt_handle* h, const clt_lma_params* prm, double* extrinsic_out, float* disparity_out,
                       double* history, int32_t* iterations) {
    if (!h || !prm || prm->num_iter < 0) return CLT_E_ARG;
    int rc;
    if ((rc = geo_prepare(h, "clt_lma_extrinsics"))) return rc;
    const int n = h->num_tasks;
    double lambda = prm->lambda;
    clt_port_geometry cur = h->pgeom, cand = h->pgeom;
    float* d_cur = h->d_gdisp;
    float* d_new = h->d_gdisp2;
    double sums[GEO_ADJ_SUMS];
    tbegin(h, T_GEOADJ);
    if ((rc = geo_eval(h, d_cur, lambda, sums))) return rc;
    double loss = sums[104];
    if (history) history[0] = loss;
    int it = 0, hist_n = 1;
    // coarse stage: Gauss-Newton steps that cancel the pairs' correlation-peak offsets
    const int blocks = (n + GEO_ADJ_TASKS - 1) / GEO_ADJ_TASKS;
    for (; it < std::min(prm->coarse_iter, prm->num_iter); it++) {
        hipLaunchKernelGGL(k_port_centres, dim3((4 * n + 255) / 256), dim3(256), 0, h->stream, h->d_port_geom, h->d_rtab,
                           h->d_gwords, h->d_gtxy, d_cur, n, h->task_words, h->d_tasks, nullptr, 0, 1.0);
        launch_convert(h, n, nullptr, 0);
        launch_corr(h, n, nullptr);
        hipLaunchKernelGGL(k_corr_offsets, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->d_corr, h->d_tasks,
                           h->task_words, n, (float)prm->coarse_radius, std::max((float)prm->coarse_strength_floor, h->cp.min_corr),
                           h->d_adj, nullptr);
        hipLaunchKernelGGL(k_geometry_adjoint, dim3(blocks), dim3(128), 0, h->stream, h->d_port_geom, h->d_rtab,
                           h->d_gwords, h->d_gtxy, d_cur, h->d_adj, n, 0.0, h->d_gpart, h->d_tstate);
        hipLaunchKernelGGL(k_sum_rows, dim3(GEO_ADJ_SUMS), dim3(256), 0, h->stream, h->d_gpart, blocks, GEO_ADJ_SUMS,
                           h->d_gsum);
        HIPCHK(h, hipGetLastError());
        double osums[GEO_ADJ_SUMS], de[13];
        HIPCHK(h, hipMemcpyAsync(osums, h->d_gsum, sizeof(osums), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, sync_streams(h));
        if (!solve13(osums, 1e-3, de)) break;
        for (int i = 0; i < 13; i++) cand.extrinsic[i] = cur.extrinsic[i] + de[i];
        HIPCHK(h, hipMemcpyAsync(h->d_gsum + GEO_ADJ_SUMS, de, sizeof(de), hipMemcpyHostToDevice, h->stream));
        if (prm->adjust_disparity)
            hipLaunchKernelGGL(k_disparity_step, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->d_tstate,
                               h->d_gsum + GEO_ADJ_SUMS, d_cur, d_new, n, 7.0);
        else
            HIPCHK(h, hipMemcpyAsync(d_new, d_cur, (size_t)n * 4, hipMemcpyDeviceToDevice, h->stream));
        cur = cand;
        std::swap(d_cur, d_new);
        // the loss and the LMA system at the new point (for the history and the next step)
        if ((rc = upload_geometry(h, cur)) || (rc = geo_eval(h, d_cur, lambda, sums))) return rc;
        loss = sums[104];
        if (history) history[hist_n++] = loss;
    }
    for (; it < prm->num_iter; it++) {
        double de[13];
        while (!solve13(sums, lambda, de)) {   // make the damped system positive definite
            lambda *= prm->lambda_scale_bad;
            if (lambda > prm->lambda_max) break;
            if ((rc = geo_eval(h, d_cur, lambda, sums))) return rc;
        }
        if (lambda > prm->lambda_max) break;
        for (int i = 0; i < 13; i++) cand.extrinsic[i] = cur.extrinsic[i] + de[i];
        HIPCHK(h, hipMemcpyAsync(h->d_gsum + GEO_ADJ_SUMS, de, sizeof(de), hipMemcpyHostToDevice, h->stream));
        if (prm->adjust_disparity)
            hipLaunchKernelGGL(k_disparity_step, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->d_tstate,
                               h->d_gsum + GEO_ADJ_SUMS, d_cur, d_new, n, 0.5);
        else
            HIPCHK(h, hipMemcpyAsync(d_new, d_cur, (size_t)n * 4, hipMemcpyDeviceToDevice, h->stream));
        if ((rc = upload_geometry(h, cand)) || (rc = geo_eval(h, d_new, lambda * prm->lambda_scale_good, sums))) return rc;
        const double new_loss = sums[104];
        if (new_loss < loss) {   // accept: the system just computed is the next one
            const double gain = (loss - new_loss) / std::max(std::fabs(loss), 1e-300);
            cur = cand;
            std::swap(d_cur, d_new);
            loss = new_loss;
            lambda *= prm->lambda_scale_good;
            if (history) history[hist_n++] = loss;
            if (gain < prm->rms_diff) { it++; break; }
        } else {                 // reject: back to the current point with more damping
            lambda *= prm->lambda_scale_bad;
            if (lambda > prm->lambda_max) { it++; break; }
            if ((rc = upload_geometry(h, cur)) || (rc = geo_eval(h, d_cur, lambda, sums))) return rc;
            if (history) history[hist_n++] = loss;
        }
    }
    tend(h, T_GEOADJ);
    // leave the handle at the solution: geometry, disparities and task records
    if ((rc = upload_geometry(h, cur)) || (rc = geo_eval(h, d_cur, 0.0, nullptr))) return rc;
    if (d_cur != h->d_gdisp) HIPCHK(h, hipMemcpyAsync(h->d_gdisp, d_cur, (size_t)n * 4, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, sync_streams(h));
    h->pgeom = cur;
    h->converted = h->corr_done = h->tex_done = h->combined = false;
    h->adj_done = true;
    if (extrinsic_out)
        for (int i = 0; i < 13; i++) extrinsic_out[i] = cur.extrinsic[i];
    if (disparity_out && n > 0) HIPCHK(h, hipMemcpy(disparity_out, h->d_gdisp, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (iterations) *iterations = hist_n - 1;
    return CLT_OK;
}

// ------------------------------------------------------------------------ batched frames
int clt_set_image_frame(clt_handle* h, int32_t frame, int32_t cam, const float* image, int32_t stride) {
    if (!h) return CLT_E_ARG;
    const clt_geom& g = h->g;
    if (frame < 0 || frame >= h->num_frames || cam < 0 || cam >= g.num_cams || !image || stride < g.width)
        return fail(h, CLT_E_ARG, "clt_set_image_frame: bad frame %d / camera %d, null image or stride %d", frame, cam, stride);
    HIPCHK(h, sync_streams(h));
    HIPCHK(h, hipMemcpy2D(h->d_images + frame * h->img_fstride + (size_t)cam * g.width * g.height, (size_t)g.width * 4, image,
                          (size_t)stride * 4, (size_t)g.width * 4, g.height, hipMemcpyHostToDevice));
    if (frame == 0) h->images_set[cam] = true;
    return CLT_OK;
}

int clt_set_tasks_frame(clt_handle* h, int32_t frame, const void* tasks, int32_t num_tasks) {
    if (!h) return CLT_E_ARG;
    if (frame < 0 || frame >= h->num_frames || num_tasks < 0 || num_tasks > h->max_tasks || (num_tasks > 0 && !tasks))
        return fail(h, CLT_E_ARG, "clt_set_tasks_frame: frame %d, %d tasks (capacity %d)", frame, num_tasks, h->max_tasks);
    const int32_t* t = (const int32_t*)tasks;
    std::vector<char> seen((size_t)h->num_tiles, 0);
    bool fusable = true;
    for (int i = 0; i < num_tasks; i++) {
        const int32_t* r = t + (size_t)i * h->task_words;
        const int tx = r[1] & 0xffff, ty = (r[1] >> 16) & 0xffff;
        if (tx >= h->tiles_x || ty >= h->tiles_y)
            return fail(h, CLT_E_ARG, "clt_set_tasks_frame: task %d tile (%d,%d) outside %dx%d", i, tx, ty, h->tiles_x, h->tiles_y);
        char& m = seen[(size_t)ty * h->tiles_x + tx];
        fusable = fusable && r[0] != 0 && !m;
        m = 1;
    }
    HIPCHK(h, sync_streams(h));   // the staging area may still be read
    h->frame_fusable[frame] = fusable;
    if (num_tasks > 0)
        HIPCHK(h, hipMemcpy(h->d_staged + (size_t)frame * h->max_tasks * h->task_words, tasks,
                            (size_t)num_tasks * h->task_words * 4, hipMemcpyHostToDevice));
    h->frame_count[frame] = num_tasks;
    h->batch_dirty = true;
    return CLT_OK;
}

// Flat task list of all frames, the frame of each task and the per-frame tile maps, packed on the
// device from the staged per-frame lists.
static int assemble_batch(clt_handle* h) {
    if (!h->batch_dirty) return CLT_OK;
    const int F = h->num_frames;
    if (F == 1) return fail(h, CLT_E_UNSUPPORTED, "clt_exec_batch: needs a handle of clt_create_batch");
    std::vector<int32_t> cs(2 * F);
    int total = 0;
    for (int f = 0; f < F; f++) {
        h->frame_start[f] = total;
        cs[f] = h->frame_count[f];
        cs[F + f] = total;
        total += h->frame_count[f];
    }
    h->frame_start[F] = total;
    HIPCHK(h, hipMemcpyAsync(h->d_fcounts, cs.data(), cs.size() * 4, hipMemcpyHostToDevice, h->stream));
    const long long map_n = (long long)F * h->num_tiles;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)std::min<long long>((map_n + 255) / 256, 4096)), dim3(256), 0, h->stream,
                       h->d_tile_map, map_n, -1);
    const long long slots = (long long)F * h->max_tasks;
    hipLaunchKernelGGL(k_batch_assemble, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, h->stream, h->d_staged,
                       h->d_fcounts, h->d_fcounts + F, F, h->max_tasks, h->task_words, h->d_tasks, h->d_task_frame,
                       h->d_tile_map, h->gd);
    HIPCHK(h, hipGetLastError());
    h->num_tasks = total;
    h->batch_dirty = false;
    return CLT_OK;
}

enum { STAGE_CONVERT = 1, STAGE_CORR = 2, STAGE_TEX = 4, STAGE_IMCLT = 8 };

// The batch as a software pipeline over tune_chunks runs of whole frames.  Each chunk runs convert,
// textures and correlate2D on the main stream (in the quad layout convert and correlate2D as one
// kernel, before the textures) and forks its imclt_rbg (after convert) and its per-tile LMA (after
// correlate2D) onto their own streams, where they overlap the next chunk's kernels; when every frame's tasks own distinct tiles with non-zero flags, textures and
// imclt_rbg of the chunk run as one pass (k_tex_imclt over the chunk's frames) instead.  Chunks
// touch disjoint frames of the CLT, tile maps and image planes and disjoint task records, so the
// only joins are the fork points and the end of the batch.
static void launch_batch(clt_handle* h, int stages) {
    const int32_t* tf = h->num_frames > 1 ? h->d_task_frame : nullptr;
    const int F = h->num_frames;
    const int chunks = std::max(1, std::min(h->tune_chunks, F));
    const bool fused = h->batch_fused && (stages & STAGE_TEX) && (stages & STAGE_IMCLT);
    // convert and correlate2D of a chunk in one kernel (quad layout), its LMA forked before the textures
    const bool conv_corr = (stages & STAGE_CONVERT) && (stages & STAGE_CORR) && h->corr_quad;
    join_side(h);   // a running imclt_rbg still reads the CLT
    join_lma(h);    // ... and an LMA the correlation records
    join_lma2(h);
    for (int c = 0; c < chunks; c++) {
        const int f0 = (int)((long)F * c / chunks), f1 = (int)((long)F * (c + 1) / chunks);
        const int t0 = F > 1 ? h->frame_start[f0] : 0;
        const int n = (F > 1 ? h->frame_start[f1] : h->num_tasks) - t0;
        if (conv_corr && n > 0) {
            launch_convert_corr_range(h, t0, n, tf, 0);
            launch_corr_lma(h, n, false, t0);
        } else if ((stages & STAGE_CONVERT) && n > 0) {
            launch_convert_range(h, t0, n, tf, 0);
        }
        if (fused) {   // textures and imclt_rbg of the chunk's frames in one pass
            launch_tex_imclt(h, f0, f1 - f0);
        } else if (stages & STAGE_IMCLT) {
            (void)hipEventRecord(h->ev_side_fork, h->stream);
            (void)hipStreamWaitEvent(h->stream_side, h->ev_side_fork, 0);
            launch_imclt(h, f1 - f0, 1, h->stream_side, f0);
            (void)hipEventRecord(h->ev_side, h->stream_side);
            h->side_pending = true;
        }
        if (!fused && (stages & STAGE_TEX) && n > 0) launch_tex(h, n, tf, nullptr, t0);
        if (!conv_corr && (stages & STAGE_CORR) && n > 0) {
            launch_corr_range(h, t0, n, tf);
            launch_corr_lma(h, n, false, t0);
        }
    }
    join_lma(h);    // a captured graph must end on the origin stream
    join_side(h);
}

static int prepare_batch(clt_handle* h, int stages) {
    for (int c = 0; c < h->g.num_cams; c++)
        if (!h->images_set[c]) return fail(h, CLT_E_STATE, "clt_exec_batch: image of camera %d of frame 0 not set", c);
    int rc;
    if ((rc = assemble_batch(h))) return rc;
    if (h->num_tasks <= 0) return fail(h, CLT_E_STATE, "clt_exec_batch: no tasks");
    if ((stages & STAGE_CORR) && (rc = prepare_corr(h))) return rc;
    if ((stages & STAGE_TEX) && (rc = prepare_tex(h))) return rc;
    const clt_geom& g = h->g;
    bool fused = g.num_colors == 3 && g.num_cams >= 2 && g.num_cams <= 5;
    for (int f = 0; f < h->num_frames; f++) fused = fused && h->frame_fusable[f];
    if (fused != h->batch_fused) h->graph_epoch++;   // the captured launches differ
    h->batch_fused = fused;
    if ((stages & STAGE_IMCLT) &&
        (rc = dalloc(h, (void**)&h->d_rbg, (size_t)h->num_frames * g.num_cams * g.num_colors * (g.width + 8) * (g.height + 8) * 4)))
        return rc;
    return CLT_OK;
}

// Time one stage over the whole batch (median of 3 after one warm-up), ms.
static float time_stage(clt_handle* h, int stage) {
    hipEvent_t b, e;
    if (hipEventCreate(&b) != hipSuccess) return 1e30f;
    if (hipEventCreate(&e) != hipSuccess) {
        (void)hipEventDestroy(b);
        return 1e30f;
    }
    float t[3] = {1e30f, 1e30f, 1e30f};
    launch_batch(h, stage);
    for (int i = 0; i < 3; i++) {
        (void)hipEventRecord(b, h->stream);
        launch_batch(h, stage);
        (void)hipEventRecord(e, h->stream);
        if (hipEventSynchronize(e) != hipSuccess || hipEventElapsedTime(&t[i], b, e) != hipSuccess) t[i] = 1e30f;
    }
    (void)hipEventDestroy(b);
    (void)hipEventDestroy(e);
    std::sort(t, t + 3);
    return t[1];
}

int clt_autotune(clt_handle* h) {
    if (!h) return CLT_E_ARG;
    int rc;
    if ((rc = prepare_batch(h, STAGE_CONVERT | STAGE_CORR | STAGE_IMCLT))) return rc;
    launch_batch(h, STAGE_CONVERT);   // every candidate below reads this CLT
    if (h->corr_quad) {   // tiles (wavefronts) per workgroup of correlate2D, and so its LDS per workgroup
        const int cand[3] = {1, 2, 4};
        float best = 1e30f;
        int pick = h->tune_corr_tpb;
        for (int c : cand) {
            h->tune_corr_tpb = c;
            const float ms = time_stage(h, STAGE_CORR);
            if (ms < best) { best = ms; pick = c; }
        }
        h->tune_corr_tpb = pick;
    }
    {   // tile rows per imclt segment (LDS carry reuse against the number of workgroups)
        const int cand[3] = {8, 16, 32};
        float best = 1e30f;
        int pick = h->tune_imclt_rows;
        for (int c : cand) {
            h->tune_imclt_rows = c;
            const float ms = time_stage(h, STAGE_IMCLT);
            if (ms < best) { best = ms; pick = c; }
        }
        h->tune_imclt_rows = pick;
    }
    int stages = STAGE_CONVERT | STAGE_CORR | STAGE_IMCLT;
    if (h->g.num_colors == 3 && h->g.num_cams >= 2 && prepare_tex(h) == CLT_OK) stages |= STAGE_TEX;
    h->err.clear();
    if ((stages & STAGE_TEX) && h->batch_fused) {   // the one-pass textures/imclt: tile rows per segment
        // (the halo re-read, 1/rows, against the workgroups) and the LDS tile layout (padded rows or
        // swizzled column groups: LDS per workgroup against the address arithmetic)
        const int cand[3] = {4, 8, 16};
        float best = 1e30f;
        int pick = h->tex_imclt_rows, pick_lay = h->tex_lds_layout;
        for (int lay = 0; lay < 2; lay++)
            for (int c : cand) {
                h->tex_imclt_rows = c;
                h->tex_lds_layout = lay;
                const float ms = time_stage(h, STAGE_TEX | STAGE_IMCLT);
                if (ms < best) { best = ms; pick = c; pick_lay = lay; }
            }
        h->tex_imclt_rows = pick;
        h->tex_lds_layout = pick_lay;
    }
    {   // frame chunks of the pipelined batch (overlap of LMA and imclt_rbg with the next chunk)
        const int cand[4] = {1, 2, 4, 8};
        float best = 1e30f;
        int pick = h->tune_chunks;
        for (int c : cand) {
            if (c > h->num_frames) break;
            h->tune_chunks = c;
            const float ms = time_stage(h, stages);
            if (ms < best) { best = ms; pick = c; }
        }
        h->tune_chunks = pick;
    }
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, sync_streams(h));
    h->tuned = true;
    h->graph_stages = 0;
    h->converted = h->corr_done = h->tex_done = false;
    return CLT_OK;
}

int clt_get_tune(clt_handle* h, int32_t* corr_tiles_per_block, int32_t* imclt_rows) {
    if (!h) return CLT_E_ARG;
    if (corr_tiles_per_block) *corr_tiles_per_block = h->tune_corr_tpb;
    if (imclt_rows) *imclt_rows = h->tune_imclt_rows;
    return CLT_OK;
}

int clt_set_batch_chunks(clt_handle* h, int32_t chunks) {
    if (!h) return CLT_E_ARG;
    if (chunks < 1 || chunks > 64) return fail(h, CLT_E_ARG, "clt_set_batch_chunks: %d", chunks);
    h->tune_chunks = chunks;
    h->graph_epoch++;
    return CLT_OK;
}

int clt_get_tune_all(clt_handle* h, int32_t* values, int32_t n) {
    if (!h || !values || n < 0) return CLT_E_ARG;
    const int32_t v[5] = {h->tune_corr_tpb, h->tune_imclt_rows, std::max(1, std::min(h->tune_chunks, h->num_frames)),
                          h->tex_imclt_rows, h->tex_lds_layout};
    for (int i = 0; i < n && i < 5; i++) values[i] = v[i];
    return CLT_OK;
}

int clt_set_tune_all(clt_handle* h, const int32_t* values, int32_t n) {
    if (!h || (!values && n > 0) || n < 0 || n > 5) return CLT_E_ARG;
    const int32_t lo[5] = {1, 1, 1, 1, 0}, hi[5] = {4, 64, 64, 64, 1};
    for (int i = 0; i < n; i++)
        if (values[i] < lo[i] || values[i] > hi[i] || (i == 0 && values[i] == 3))
            return fail(h, CLT_E_ARG, "clt_set_tune_all: value %d = %d", i, values[i]);
    if (n > 0) h->tune_corr_tpb = values[0];
    if (n > 1) h->tune_imclt_rows = values[1];
    if (n > 2) h->tune_chunks = values[2];
    if (n > 3) h->tex_imclt_rows = values[3];
    if (n > 4) h->tex_lds_layout = values[4];
    h->graph_epoch++;
    return CLT_OK;
}

int clt_get_batch_chunks(clt_handle* h, int32_t* chunks) {
    if (!h || !chunks) return CLT_E_ARG;
    *chunks = std::max(1, std::min(h->tune_chunks, h->num_frames));
    return CLT_OK;
}

int clt_exec_batch(clt_handle* h, int32_t stages) {
    if (!h) return CLT_E_ARG;
    if (stages <= 0 || stages > 15 || ((stages & ~STAGE_CONVERT) && !(stages & STAGE_CONVERT) && !h->converted))
        return fail(h, CLT_E_ARG, "clt_exec_batch: stages 0x%x", stages);
    int rc;
    if ((rc = prepare_batch(h, stages))) return rc;
    if (h->graph_stages != stages || h->graph_tasks != h->num_tasks ||   // grid sizes follow the task count
        h->graph_epoch_cap != h->graph_epoch) {                             // launch arguments changed
        if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
        if (h->graph) (void)hipGraphDestroy(h->graph);
        h->graph_exec = nullptr;
        h->graph = nullptr;
        HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        launch_batch(h, stages);
        HIPCHK(h, hipStreamEndCapture(h->stream, &h->graph));
        HIPCHK(h, hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
        h->graph_stages = stages;
        h->graph_tasks = h->num_tasks;
        h->graph_epoch_cap = h->graph_epoch;
    }
    tbegin(h, T_BATCH);
    HIPCHK(h, hipGraphLaunch(h->graph_exec, h->stream));
    tend(h, T_BATCH);
    if (stages & STAGE_CONVERT) h->converted = true;
    if (stages & STAGE_CORR) h->corr_done = true;
    if (stages & STAGE_TEX) h->tex_done = true;
    return finish(h);
}

static int frame_range(clt_handle* h, int frame, size_t* first, size_t* count) {
    if (frame < 0 || frame >= h->num_frames) return fail(h, CLT_E_ARG, "frame %d of %d", frame, h->num_frames);
    if (h->num_frames == 1) {
        *first = 0;
        *count = h->num_tasks;
    } else {
        if (h->batch_dirty) return fail(h, CLT_E_STATE, "tasks changed since the last clt_exec_batch");
        *first = h->frame_start[frame];
        *count = h->frame_start[frame + 1] - h->frame_start[frame];
    }
    return CLT_OK;
}

int clt_get_corr2d_frame(clt_handle* h, int32_t frame, float* corr, float* combo, float* disparity) {
    if (!h) return CLT_E_ARG;
    if (!h->corr_done) return fail(h, CLT_E_STATE, "clt_get_corr2d_frame: corr2d not executed");
    size_t f0, nt;
    int rc;
    if ((rc = frame_range(h, frame, &f0, &nt))) return rc;
    HIPCHK(h, sync_streams(h));
    const size_t P = h->corr_pairs_at_exec, sh = h->corr_strip_h_at_exec;
    if (corr && nt) HIPCHK(h, hipMemcpy(corr, h->d_corr + f0 * P * CLT_CORR_LEN, nt * P * CLT_CORR_LEN * 4, hipMemcpyDeviceToHost));
    if (combo && sh && nt) HIPCHK(h, hipMemcpy(combo, h->d_combo + f0 * sh * 15, nt * sh * 15 * 4, hipMemcpyDeviceToHost));
    if (disparity && sh && nt)
        HIPCHK(h, hipMemcpy(disparity, h->d_disp + f0 * CLT_DISP_LAYERS, nt * CLT_DISP_LAYERS * 4, hipMemcpyDeviceToHost));
    return CLT_OK;
}

int clt_get_textures_frame(clt_handle* h, int32_t frame, float* tiles, float* stats) {
    if (!h) return CLT_E_ARG;
    if (!h->tex_done) return fail(h, CLT_E_STATE, "clt_get_textures_frame: textures not executed");
    size_t f0, nt;
    int rc;
    if ((rc = frame_range(h, frame, &f0, &nt))) return rc;
    HIPCHK(h, sync_streams(h));
    const size_t pl = h->tex_planes_at_exec, C = h->g.num_cams;
    if (tiles && nt) HIPCHK(h, hipMemcpy(tiles, h->d_tex + f0 * pl * 256, nt * pl * 256 * 4, hipMemcpyDeviceToHost));
    if (stats && nt) HIPCHK(h, hipMemcpy(stats, h->d_tex_stats + f0 * 4 * C, nt * 4 * C * 4, hipMemcpyDeviceToHost));
    return CLT_OK;
}

int clt_get_rbg_frame(clt_handle* h, int32_t frame, int32_t cam, float* out) {
    if (!h || !out) return CLT_E_ARG;
    if (frame < 0 || frame >= h->num_frames || cam < 0 || cam >= h->g.num_cams)
        return fail(h, CLT_E_ARG, "clt_get_rbg_frame: frame %d / camera %d", frame, cam);
    if (!h->d_rbg) return fail(h, CLT_E_STATE, "clt_get_rbg_frame: imclt not executed");
    const size_t n = (size_t)h->g.num_colors * (h->g.width + 8) * (h->g.height + 8);
    HIPCHK(h, sync_streams(h));
    HIPCHK(h, hipMemcpy(out, h->d_rbg + ((size_t)frame * h->g.num_cams + cam) * n, n * 4, hipMemcpyDeviceToHost));
    return CLT_OK;
}

// ------------------------------------------------------------------------ control / timing
int clt_set_async(clt_handle* h, int32_t async) {
    if (!h) return CLT_E_ARG;
    h->async = async != 0;
    return CLT_OK;
}

int clt_sync(clt_handle* h) {
    if (!h) return CLT_E_ARG;
    HIPCHK(h, sync_streams(h));
    return CLT_OK;
}

int clt_set_timing(clt_handle* h, int32_t enable) {
    if (!h || enable < 0) return CLT_E_ARG;
    h->timing = enable != 0;
    for (int i = 0; i < T_NUM && enable > 1; i++)   // events for `enable` launches of each kernel, created now
        while ((int)h->ev_beg[i].size() < enable) {
            hipEvent_t b = nullptr, e = nullptr;   // timing only, as in tbegin
            HIPCHK(h, hipEventCreateWithFlags(&b, hipEventDisableSystemFence));
            if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
                (void)hipEventDestroy(b);
                return fail(h, CLT_E_HIP, "clt_set_timing: hipEventCreate failed");
            }
            h->ev_beg[i].push_back(b);
            h->ev_end[i].push_back(e);
        }
    return CLT_OK;
}

int clt_get_timing(clt_handle* h, int32_t which, double* total_ms, int32_t* count) {
    if (!h || which < 0 || which >= T_NUM) return CLT_E_ARG;
    HIPCHK(h, sync_streams(h));
    double tot = 0;
    for (int u = 0; u < h->ev_used[which]; u++) {
        float ms = 0.f;
        HIPCHK(h, hipEventElapsedTime(&ms, h->ev_beg[which][u], h->ev_end[which][u]));
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (count) *count = h->ev_used[which];
    return CLT_OK;
}

int clt_reset_timing(clt_handle* h) {
    if (!h) return CLT_E_ARG;
    HIPCHK(h, sync_streams(h));
    for (int i = 0; i < T_NUM; i++) h->ev_used[i] = 0;
    return CLT_OK;
}

}  // extern "C"
