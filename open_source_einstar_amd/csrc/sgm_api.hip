// C-ABI layer of the gfx950 SGM library (include/sgm_hip.h): contexts, the per-pair pipeline,
// and batches. Kernels live in sgm_kernels.hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "sgm_hip.h"
#include "sgm_internal.h"

using namespace sgm;

namespace {

thread_local std::string g_last_error;

// ev[0..4] bracket census, paths, rows and median; ev[5] marks the start of the front end
// (rectification and pyramid) when the call had one (front); ev[6..8] bracket the refinement and the
// filters of compute_disparity (post); ev[9] ends the point image of a frame call (points).
constexpr int kStageEvents = 10;
struct Stages {
    hipEvent_t ev[kStageEvents] = {};
    bool recorded = false;
    bool front = false;
    bool post = false;
    bool points = false;
    int last() const { return points ? 9 : (post ? 8 : 4); }
};

// Stage events of the calls in flight: a ring, so that per-stage times can be summed over many
// asynchronous calls (bench loops) and not only read for the last one.
constexpr int kEventRing = 64;

std::string hip_msg(const char* what, hipError_t e) { return std::string(what) + ": " + hipGetErrorString(e); }

// Parameter validation of MetalStereo::create (metal_stereo.cpp:115-120), for byte path costs.
int validate_params(const sgm_params* p, std::string& err) {
    if (!p) {
        err = "params is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (p->num_disparities < 1 || p->num_disparities > 512) {
        err = "num_disparities must be in [1, 512]";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (p->census_radius_x < 0 || p->census_radius_y < 0 || 2 * p->census_radius_x + 1 > 9 || 2 * p->census_radius_y + 1 > 7) {
        err = "census window too large for 64 bits (at most 9x7)";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (p->p1 < 0 || p->p1 > (1 << 30) || p->p2 < 0) {
        err = "p1 must be in [0, 2^30] and p2 >= 0";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    // MetalStereo::create's check and message (metal_stereo.cpp:118-120). Path costs are bytes where
    // the largest step penalty allows it and u16 otherwise (wide_path_costs).
    const int cells = (2 * p->census_radius_x + 1) * (2 * p->census_radius_y + 1);
    if (cells + p->p2 > 255) {
        err = "census window + P2 exceed the byte-sized SGM costs";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    return SGM_OK;
}

// Rows wider than kMaxStagedWidth need 4 bytes of LDS per column in one workgroup (the row kernel's
// right WTA keys, the occlusion filter's claims): refused where the device cannot give that much.
int check_row_lds(int w, std::string& err) {
    if (w <= kMaxStagedWidth) return SGM_OK;
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
        err = "cannot query the device's LDS size for a row of " + std::to_string(w) + " columns";
        return SGM_ERR_IO;
    }
    if ((int64_t)w * 4 + 16 > lds) {
        err = "rows of " + std::to_string(w) + " columns need " + std::to_string((int64_t)w * 4 + 16) +
              " bytes of LDS per workgroup; this device has " + std::to_string(lds);
        return SGM_ERR_UNSUPPORTED;
    }
    return SGM_OK;
}

SgmConsts make_consts(const sgm_params& p) {
    SgmConsts c{};
    c.p1 = p.p1;
    c.p2 = p.p2;
    c.adaptive_p2 = p.adaptive_p2 ? 1 : 0;
    c.slant_steps = p.slant_steps ? 1 : 0;
    c.wide_costs = wide_path_costs((2 * p.census_radius_x + 1) * (2 * p.census_radius_y + 1), p.p1, p.p2, p.adaptive_p2 != 0) ? 1 : 0;
    c.ndirs = p.eight_paths ? 8 : 4;
    c.census_rx = p.census_radius_x;
    c.census_ry = p.census_radius_y;
    c.invalid_cost = (2 * p.census_radius_x + 1) * (2 * p.census_radius_y + 1);  // stereo.cpp:366
    c.lr_max_diff = p.lr_max_diff;
    c.subpixel = p.subpixel ? 1 : 0;
    c.uniqueness = p.uniqueness;
    // stereo.cpp:64: p2 = max(P1+1, P2 / (|dI|/8 + 1)); |dI| <= 255 so |dI|/8 <= 31.
    for (int q = 0; q < 32; ++q) c.p2_by_q[q] = std::max(p.p1 + 1, p.p2 / (q + 1));
    return c;
}

template <typename T>
hipError_t dalloc(T** ptr, size_t count, size_t& total) {
    *ptr = nullptr;
    if (count == 0) return hipSuccess;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(ptr), count * sizeof(T));
    if (e == hipSuccess) total += count * sizeof(T);
    return e;
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

// Lanes per path scanline when the caller leaves it open (0). Wider lines cut each lane's share of a
// step (the walk's serial chain) and multiply the lanes a scanline occupies, so they pay where the
// walks are few for the card or long per step. 64 lanes (4 or 8 disparities per lane): layouts of 32
// per 16-lane lane (D > 384; path kernel 0.54-0.87 of the 16-lane time from 320x256 to 2048x1536),
// and layouts of 16 when half the directions or more recompute their costs (0.45-0.76; 0.77-1.04 of
// the 32-lane time). 32 lanes: other layouts of a multiple of 8 per lane that recompute (0.42-0.89),
// 4-path pairs from the cost volume at 16 per lane up to 2M pixels (D = 256: 0.91-0.97 of the best
// 16-lane form from 320x256 to 1600x1200), at 8 per lane up to 128K pixels (320x256x128: 40 against
// 46 us; larger pairs run 16 lanes of 16 lines faster) and at 24 or more per lane up to 1M pixels, and
// 8-path pairs of 16 or more per lane up to 640x512 (0.78-0.93). With 2 or 6 per lane a lane would
// move its path costs in 2-byte pieces (D = 192, 12 per lane, on 32 lanes: 1.9-2.4 times the 16-lane
// time), so wider lines need a multiple of 8 per 16-lane lane (32 lanes) or of 16 (64 lanes).
static bool mostly_recomputed(const PairDesc& d, const SgmConsts& c) {
    const int recomputed = d.cvol ? std::min(std::max(c.cv_from_dir, 0), d.ndirs) : d.ndirs;
    return 2 * recomputed >= d.ndirs;
}

int auto_path_group(const PairDesc& d, const SgmConsts& c) {
    if (c.wide_costs || d.dpl % 8 != 0) return 16;
    const int64_t px = (int64_t)d.w * d.h;
    const bool recomputing = mostly_recomputed(d, c);
    if (d.dpl % 16 == 0 && (d.dpl >= 32 || recomputing)) return 64;
    if (recomputing) return 32;
    if (d.ndirs == 4) {
        const int64_t limit = d.dpl == 8 ? 128 * 1024 : d.dpl == 16 ? 2048 * 1024 : 1024 * 1024;
        return px <= limit ? 32 : 16;
    }
    return d.dpl >= 16 && px <= 640 * 512 ? 32 : 16;
}

// Scanlines per path-kernel block when the caller leaves it open (0). 16 lanes per line, 4 paths: one
// wave per block (4 lines) for pairs up to 640x400 and for pairs that recompute most directions' costs
// or hold more than 16 disparities per lane (1280x1024x128 recomputing: 0.55 against 0.63 ms); 16
// lines for larger cost-volume pairs of up to 16 per lane (0.68-0.97 of the 4-line time from 640x480
// to 2048x1536 at D = 64 to 192, except 1.01-1.05 at 1280x1024 and 1600x1200x192). 8 paths: 16 lines
// (1280x1024: 0.47-0.49 ms either way; 2048x1536x256: 2.57 against 2.64 ms; the 320x256x32 pair of the
// refinement chain: 33 against 41 us). 32 lanes per line: 8 lines (two waves), except 4 for 4-path pairs of 24 or more
// disparities per lane of the 16-lane layout (640x512, D = 384 and 512: 198 and 248 against 214 and
// 284 us). 64 lanes: 4 lines (one a wave).
int auto_path_lines(const PairDesc& d, const SgmConsts& c, int path_group) {
    if (path_group == 64) return 4;
    if (path_group == 32) return d.ndirs == 4 && d.dpl >= 24 ? 4 : 8;
    if (d.ndirs != 4) return 16;
    const bool long_lines =
        (int64_t)d.w * d.h > 640 * 400 && d.dpl <= 16 && !c.wide_costs && !mostly_recomputed(d, c);
    return long_lines ? 16 : 4;
}

// Threads per row-kernel block when the caller leaves it open (0): the large block (1024 threads, or
// 512) for pairs of at most kBigRowBlockRows rows -- one row per CU of the MI355X, so the large blocks
// still run in one round -- where 256-thread blocks would leave most wave slots empty and give each
// 16-lane group a long serial run of pixels (row kernel 0.53-0.86 of the 256-thread time from 160x120
// to 4096x200); 256 otherwise (384 rows and more: 1.01-1.08 times the 256-thread time).
constexpr int kBigRowBlockRows = 256;
int auto_row_threads(const PairDesc& d) { return d.h <= kBigRowBlockRows ? 1024 : kBlock; }

// Enqueues the stages of one pair (or one launch table). Records stage events when ev != null.
hipError_t enqueue_pair(const PairDesc& d, const SgmConsts& c, bool padded, hipStream_t s, Stages* st, int path_lines = 0,
                        int path_group = 0, int row_threads = 0) {
    if (row_threads <= 0) row_threads = auto_row_threads(d);
    if (path_group <= 0) path_group = auto_path_group(d, c);
    if (path_lines <= 0) path_lines = auto_path_lines(d, c, path_group);
    // A 64-lane line needs dpl % 16 == 0 and 4 lines per block, else it runs 32 lanes; a 32-lane line
    // needs dpl % 4 == 0 and at most 8 lines per block, else it runs 16.
    if (path_group == 64 && !paths_group_ok(d.dpl, 64, path_lines)) {
        if (d.dpl % 16 == 0) path_lines = 4;
        else path_group = 32;
    }
    if (path_group == 32 && !paths_group_ok(d.dpl, 32, path_lines)) {
        if (d.dpl % 4 == 0 && path_lines == 16) path_lines = 8;
        else path_group = 16;
    }
    LaunchTable t{};
    t.one = d;
    t.n = 1;
    hipError_t e;
    if (st && (e = hipEventRecord(st->ev[0], s)) != hipSuccess) return e;
    t.total_blocks = census_blocks(d.w, d.h);
    if ((e = launch_census(t, c, s)) != hipSuccess) return e;
    if (d.cost_volume) {
        t.total_blocks = (int)(((int64_t)d.w * d.h + kBlock - 1) / kBlock);
        if ((e = launch_cost_volume(t, c, s)) != hipSuccess) return e;
    }
    if (d.cvol) {
        t.total_blocks = cost8_blocks(d.w, d.h);
        if ((e = launch_cost8(d.dpl, t, c, s)) != hipSuccess) return e;
    }
    if (st && (e = hipEventRecord(st->ev[1], s)) != hipSuccess) return e;
    t.total_blocks = path_blocks(d, path_lines);
    if ((e = launch_paths(d.dpl, padded, t, c, path_lines, path_group, s)) != hipSuccess) return e;
    if (st && (e = hipEventRecord(st->ev[2], s)) != hipSuccess) return e;
    t.total_blocks = row_blocks(d);
    if ((e = launch_rows(d.dpl, padded, t, c, d.w, s, row_threads)) != hipSuccess) return e;
    if (st && (e = hipEventRecord(st->ev[3], s)) != hipSuccess) return e;
    if (d.median) {
        t.total_blocks = median_blocks(d.w, d.h);
        if ((e = launch_median(t, s)) != hipSuccess) return e;
    }
    if (st) {
        if ((e = hipEventRecord(st->ev[4], s)) != hipSuccess) return e;
        st->recorded = true;
    }
    return hipSuccess;
}

// es: bytes per stored path cost (1, or 2 for u16 path costs).
void set_geometry(PairDesc& d, int w, int h, int min_d, int nd, int ndirs, int es) {
    d.w = w;
    d.h = h;
    d.min_d = min_d;
    d.nd = nd;
    d.dpl = dpl_for_pair(nd, w, h);
    d.dp = d.dpl * kGroup;
    d.ndirs = ndirs;
    d.lvol_px_stride = (int64_t)ndirs * d.dp * es;
    census_layout(w, min_d, d.dp, d.census_padl, d.census_pitch);
    fill_lines(d);
}

// Census words of one image (padded rows) for a pair of this geometry.
size_t census_words(int w, int h, int min_d, int nd) {
    int32_t padl, pitch;
    census_layout(w, min_d, dpl_for(nd) * kGroup, padl, pitch);
    return (size_t)pitch * (size_t)h;
}

}  // namespace

namespace sgm {
// Entry points shared with the serving loop (sgm_server.hip).
int validate_params_ext(const sgm_params* p, std::string& err) { return validate_params(p, err); }
int check_row_lds_ext(int w, std::string& err) { return check_row_lds(w, err); }
SgmConsts make_consts_ext(const sgm_params& p) { return make_consts(p); }
hipError_t enqueue_pair_ext(const PairDesc& d, const SgmConsts& c, bool padded, hipStream_t s, hipEvent_t* ev5,
                            int path_group) {
    Stages st;
    if (ev5)
        for (int i = 0; i < 5; ++i) st.ev[i] = ev5[i];
    return enqueue_pair(d, c, padded, s, ev5 ? &st : nullptr, 0, path_group);
}
void set_geometry_ext(PairDesc& d, int w, int h, int min_d, int nd, int ndirs, int es) {
    set_geometry(d, w, h, min_d, nd, ndirs, es);
}
size_t census_words_ext(int w, int h, int min_d, int nd) { return census_words(w, h, min_d, nd); }
void set_thread_error(const std::string& e) { g_last_error = e; }
}  // namespace sgm

struct sgm_context {
    sgm_params p{};
    SgmConsts consts{};
    int max_w = 0, max_h = 0, ndirs = 4;
    hipStream_t stream = nullptr;
    hipStream_t last_stream = nullptr;
    hipEvent_t order_ev = nullptr;  // orders a call on another stream after the previous call's
    Stages ring[kEventRing];
    int ring_head = 0;     // next set to record
    int ring_pending = 0;  // recorded, not yet read; the oldest is ring_head - ring_pending
    sgm_stage_totals totals{};
    uint8_t* d_img = nullptr;      // left | right, compact pitch, for host inputs
    uint64_t* d_census = nullptr;  // left | right, padded rows
    size_t census_half = 0;        // words per image
    uint8_t* d_lvol = nullptr;
    float* d_disp = nullptr;
    float* d_median = nullptr;
    float* d_conf = nullptr;
    int32_t* d_dint = nullptr;
    uint8_t* d_status = nullptr;
    int32_t* d_rdisp = nullptr;
    uint16_t* d_best = nullptr;
    uint16_t* d_second = nullptr;
    uint8_t* d_cvol = nullptr;   // u8 matching costs when the cost-volume mode is on
    // SGM_OPT_VIEW_BASE_HIGH: the allocations that d_census and d_cvol were placed inside (else null).
    void* d_census_block = nullptr;
    void* d_cvol_block = nullptr;
    bool use_cvol = false;
    int path_lines = 0;          // scanlines per path-kernel block (0: auto_path_lines)
    int path_group = 0;          // lanes per scanline in the path kernel (0: auto_path_group, 16 or 32)
    int row_threads = 0;         // threads per row-kernel block (0: auto_row_threads, 256 or 1024)
    uint16_t* d_sum = nullptr;   // debug volumes, allocated on first request
    uint16_t* d_cost = nullptr;
    // Resident rectification (sgm_set_rectification): both maps as interleaved (x, y) floats, and
    // the raw frames of host calls.
    float* d_maps = nullptr;     // left | right, map_w * map_h * 2 floats each
    uint8_t* d_raw = nullptr;    // left | right, raw_w * raw_h each
    size_t maps_bytes = 0, raw_bytes = 0;
    int map_w = 0, map_h = 0, raw_w = 0, raw_h = 0;
    // compute_disparity: SGM runs at level `levels` of a pyramid over stereo_w x stereo_h images
    // (sgm_create: level 0 and no refinement or filters).
    sgm_stereo_params sp{};
    int levels = 0;
    int stereo_w = 0, stereo_h = 0;
    uint8_t* d_pyr = nullptr;    // levels 1..L, left then right of each level, compact rows
    float* d_wa = nullptr;       // level-0 sized work maps: disparity ping-pong, median, confidence
    float* d_wb = nullptr;
    float* d_wm = nullptr;
    float* d_wc = nullptr;
    int32_t* d_labels = nullptr; // speckle components: labels, tile-local counts, component sizes
    uint32_t* d_local = nullptr;
    uint32_t* d_sizes = nullptr;
    uint8_t* d_full = nullptr;   // full-resolution pair (2 * stereo_w x 2 * stereo_h each) for the detail refinement
    // Point frames (sgm_set_point_params).
    bool have_points = false;
    sgm_rectified_geometry geom{};
    sgm_point_params pparams{};
    float* d_pts = nullptr;      // float4 per pixel
    float* d_nrm = nullptr;      // float4 per pixel
    float* d_wts = nullptr;
    size_t bytes = 0;
    sgm_timings timings{};
    double pending_wall_ms = 0;
    std::string err;

    int fail(int code, std::string msg) {
        err = std::move(msg);
        return code;
    }
    int hip_fail(const char* what, hipError_t e) { return fail(SGM_ERR_IO, hip_msg(what, e)); }
};

extern "C" {

void sgm_params_default(sgm_params* p) {
    if (!p) return;
    *p = sgm_params{};
    p->min_disparity = 0;
    p->num_disparities = 64;
    p->census_radius_x = 2;
    p->census_radius_y = 2;
    p->p1 = 8;
    p->p2 = 96;
    p->adaptive_p2 = 1;
    p->eight_paths = 0;
    p->slant_steps = 1;
    p->uniqueness = 0.9f;
    p->lr_max_diff = 1;
    p->subpixel = 1;
}

const char* sgm_last_error(const sgm_context* ctx) { return ctx ? ctx->err.c_str() : g_last_error.c_str(); }

void sgm_destroy(sgm_context* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->last_stream && ctx->last_stream != ctx->stream) (void)hipStreamSynchronize(ctx->last_stream);
    for (Stages& st : ctx->ring)
        for (hipEvent_t e : st.ev)
            if (e) (void)hipEventDestroy(e);
    void* census = ctx->d_census_block ? ctx->d_census_block : (void*)ctx->d_census;
    void* cvol = ctx->d_cvol_block ? ctx->d_cvol_block : (void*)ctx->d_cvol;
    for (void* p : {(void*)ctx->d_img, census, (void*)ctx->d_lvol, (void*)ctx->d_disp, (void*)ctx->d_median,
                    (void*)ctx->d_conf, (void*)ctx->d_dint, (void*)ctx->d_status, (void*)ctx->d_rdisp, (void*)ctx->d_best,
                    (void*)ctx->d_second, (void*)ctx->d_sum, (void*)ctx->d_cost, cvol, (void*)ctx->d_maps,
                    (void*)ctx->d_raw, (void*)ctx->d_pyr, (void*)ctx->d_wa, (void*)ctx->d_wb, (void*)ctx->d_wm,
                    (void*)ctx->d_wc, (void*)ctx->d_labels, (void*)ctx->d_local, (void*)ctx->d_sizes,
                    (void*)ctx->d_full, (void*)ctx->d_pts, (void*)ctx->d_nrm, (void*)ctx->d_wts})
        dfree(p);
    if (ctx->order_ev) (void)hipEventDestroy(ctx->order_ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // extern "C"

// The SGM-only configuration sgm_create runs: compute_disparity is then sgm_disparity itself.
static sgm_stereo_params sgm_only(const sgm_params& p) {
    sgm_stereo_params sp{};
    sp.sgm = p;
    sp.pyramid_levels = 0;
    sp.refine_search_radius = 2;
    sp.refine_zncc_radius = 3;
    sp.refine_min_zncc = 0.5f;
    sp.refine_reject_edge_winners = 1;
    sp.detail_enabled = 0;
    sp.detail_radius = 5;
    sp.detail_search = 1.0f;
    sp.detail_step = 0.5f;
    sp.detail_min_zncc = 0.5f;
    sp.filter_occlusion = 0;
    sp.filter_occlusion_tolerance = 1.0f;
    sp.filter_spikes = 0;
    sp.filter_spike_threshold = 1.0f;
    sp.speckle_max_region_size = 0;
    sp.speckle_max_diff = 1.0f;
    return sp;
}

static bool has_post(const sgm_stereo_params& sp) {
    return sp.pyramid_levels > 0 || sp.detail_enabled || sp.filter_occlusion || sp.filter_spikes ||
           sp.speckle_max_region_size > 1;
}

// Limits of the refinement stages, as MetalStereo::create checks them (metal_stereo.cpp:145-149).
static int validate_stereo(const sgm_stereo_params* sp, std::string& err) {
    if (sp->pyramid_levels < 0 || sp->pyramid_levels > kMaxLevels) {
        err = "pyramid_levels must be in [0, " + std::to_string(kMaxLevels) + "]";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (sp->pyramid_levels > 0 && (sp->refine_search_radius < 1 || sp->refine_search_radius > 4)) {
        err = "refine search radius must be 1..4";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (sp->pyramid_levels > 0 && (sp->refine_zncc_radius < 1 || sp->refine_zncc_radius > 7)) {
        err = "refine zncc_radius must be 1..7";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (sp->detail_enabled && (!(sp->detail_step > 0) || std::lround(sp->detail_search / sp->detail_step) > 8 ||
                               sp->detail_radius < 1 || sp->detail_radius > 16))
        return err = "detail search must be at most 8 steps either side (step > 0, radius 1..16)", SGM_ERR_INVALID_ARGUMENT;
    return SGM_OK;
}

// A context for SGM on images up to sgm_w x sgm_h, run inside compute_disparity on stereo_w x
// stereo_h images with the stages of sp.
static int create_context(const sgm_params* params, int sgm_w, int sgm_h, int stereo_w, int stereo_h,
                          const sgm_stereo_params& sp, sgm_context** out) {
    auto* ctx = new sgm_context;
    ctx->p = *params;
    ctx->consts = make_consts(*params);
    ctx->max_w = sgm_w;
    ctx->max_h = sgm_h;
    ctx->stereo_w = stereo_w;
    ctx->stereo_h = stereo_h;
    ctx->sp = sp;
    ctx->levels = sp.pyramid_levels;
    ctx->ndirs = params->eight_paths ? 8 : 4;
    const size_t npx = (size_t)sgm_w * (size_t)sgm_h;
    const size_t spx = (size_t)stereo_w * (size_t)stereo_h;
    const size_t dp = (size_t)dpl_for(params->num_disparities) * kGroup;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->order_ev, hipEventDisableTiming);
    // Stage events only time the stages (completion is always a stream synchronisation), so they skip
    // the system-scope fence: with it, each record between two kernels cost a cache writeback, about
    // 4 us of the step.
    for (int r = 0; e == hipSuccess && r < kEventRing; ++r)
        for (int i = 0; e == hipSuccess && i < kStageEvents; ++i)
            e = hipEventCreateWithFlags(&ctx->ring[r].ev[i], hipEventDisableSystemFence);
    size_t& b = ctx->bytes;
    if (e == hipSuccess) e = dalloc(&ctx->d_img, 2 * std::max(npx, spx), b);
    ctx->census_half = census_words(sgm_w, sgm_h, params->min_disparity, params->num_disparities);
    // Census words of both images, then the left image's step-penalty words (4 bytes per pixel).
    const bool wide = ctx->consts.wide_costs != 0;
    if (e == hipSuccess) e = dalloc(&ctx->d_census, 2 * ctx->census_half + (npx * edge_bytes_per_px(wide) + 7) / 8, b);
    if (e == hipSuccess) e = dalloc(&ctx->d_lvol, (size_t)ctx->ndirs * npx * dp * (wide ? 2 : 1), b);
    // Cost-volume mode is the default wherever the volume fits one buffer view.
    ctx->use_cvol = npx * dp <= kMaxViewBytes;
    if (e == hipSuccess && ctx->use_cvol) e = dalloc(&ctx->d_cvol, npx * dp, b);
    if (e == hipSuccess) e = dalloc(&ctx->d_disp, npx, b);
    if (e == hipSuccess) e = dalloc(&ctx->d_median, npx, b);
    if (e == hipSuccess) e = dalloc(&ctx->d_conf, npx, b);
    if (e == hipSuccess) e = dalloc(&ctx->d_dint, npx, b);
    if (e == hipSuccess) e = dalloc(&ctx->d_status, npx, b);
    if (e == hipSuccess) e = dalloc(&ctx->d_rdisp, npx, b);
    if (e == hipSuccess) e = dalloc(&ctx->d_best, npx, b);
    if (e == hipSuccess) e = dalloc(&ctx->d_second, npx, b);
    if (has_post(sp)) {
        size_t pyr = 0;
        for (int i = 1, lw = stereo_w, lh = stereo_h; i <= sp.pyramid_levels; ++i) {
            lw /= 2;
            lh /= 2;
            pyr += 2 * (size_t)lw * lh;
        }
        if (e == hipSuccess && pyr) e = dalloc(&ctx->d_pyr, pyr, b);
        for (float** w : {&ctx->d_wa, &ctx->d_wb, &ctx->d_wm, &ctx->d_wc})
            if (e == hipSuccess) e = dalloc(w, spx, b);
        if (e == hipSuccess) e = dalloc(&ctx->d_labels, spx, b);
        if (e == hipSuccess) e = dalloc(&ctx->d_local, spx, b);
        if (e == hipSuccess) e = dalloc(&ctx->d_sizes, spx, b);
        if (e == hipSuccess && sp.detail_enabled) e = dalloc(&ctx->d_full, 8 * spx, b);
    }
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_create: device allocation", e);
        sgm_destroy(ctx);
        return SGM_ERR_IO;
    }
    *out = ctx;
    return SGM_OK;
}

extern "C" {

int sgm_create(const sgm_params* params, int max_width, int max_height, sgm_context** out) {
    if (!out) {
        g_last_error = "out is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    std::string err;
    if (int rc = validate_params(params, err); rc != SGM_OK) {
        g_last_error = err;
        return rc;
    }
    if (max_width < 1 || max_height < 1 || max_width > kMaxWidth) {
        g_last_error = "max_width must be in [1, " + std::to_string(kMaxWidth) + "] and max_height >= 1";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (census_words(max_width, max_height, params->min_disparity, params->num_disparities) * 8 > kMaxViewBytes) {
        g_last_error = "image too large: the census words of one image exceed 4 GiB - 1";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = check_row_lds(max_width, err); rc != SGM_OK) {
        g_last_error = err;
        return rc;
    }
    return create_context(params, max_width, max_height, max_width, max_height, sgm_only(*params), out);
}

void sgm_stereo_params_default(sgm_stereo_params* p) {
    if (!p) return;
    *p = sgm_only(sgm_params{});
    sgm_params_default(&p->sgm);
    p->pyramid_levels = 2;                 // stereo.hpp:82-89 and the stage defaults of :42-79
    p->detail_enabled = 1;
    p->filter_occlusion = 1;
    p->filter_spikes = 1;
    p->speckle_max_region_size = 200;
}

int sgm_create_stereo(const sgm_stereo_params* sp, int width, int height, sgm_context** out) {
    if (!out || !sp) {
        g_last_error = "params and out must not be NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    std::string err;
    if (int rc = validate_params(&sp->sgm, err); rc != SGM_OK) {
        g_last_error = err;
        return rc;
    }
    if (int rc = validate_stereo(sp, err); rc != SGM_OK) {
        g_last_error = err;
        return rc;
    }
    const int sw = width >> sp->pyramid_levels, sh = height >> sp->pyramid_levels;
    if (width < 1 || height < 1 || width > kMaxWidth || sw < 1 || sh < 1) {
        g_last_error = "width must be in [1, " + std::to_string(kMaxWidth) + "] and the SGM level at least 1x1";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (census_words(sw, sh, sp->sgm.min_disparity, sp->sgm.num_disparities) * 8 > kMaxViewBytes) {
        g_last_error = "image too large: the census words of one image exceed 4 GiB - 1";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if ((int64_t)width * height > INT32_MAX) {  // speckle labels are 32-bit pixel indices
        g_last_error = "image too large: more than 2^31 - 1 pixels at the full resolution";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = check_row_lds(width, err); rc != SGM_OK) {  // the occlusion filter runs at the full width
        g_last_error = err;
        return rc;
    }
    return create_context(&sp->sgm, sw, sh, width, height, *sp, out);
}

size_t sgm_device_bytes(const sgm_context* ctx) { return ctx ? ctx->bytes : 0; }

// A buffer of `bytes` inside a new allocation of bytes + 4 GiB, placed where the low 32-bit word of its
// address is `low`: *block receives the allocation, the return value the placed buffer.
static hipError_t alloc_at_low_word(size_t bytes, uint32_t low, void** block, void** placed) {
    *block = *placed = nullptr;
    hipError_t e = hipMalloc(block, bytes + (size_t(1) << 32));
    if (e != hipSuccess) return e;
    const uintptr_t base = reinterpret_cast<uintptr_t>(*block);
    *placed = reinterpret_cast<void*>(base + (uint32_t)(low - (uint32_t)base));  // distance mod 2^32
    return hipSuccess;
}

int sgm_set_option(sgm_context* ctx, int option, int64_t value) {
    if (!ctx) return SGM_ERR_INVALID_ARGUMENT;
    switch (option) {
        case SGM_OPT_VIEW_BASE_HIGH: {
            if (value != 1) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "SGM_OPT_VIEW_BASE_HIGH takes 1");
            if (ctx->d_census_block) return SGM_OK;
            hipError_t e = hipStreamSynchronize(ctx->stream);
            if (e == hipSuccess && ctx->last_stream) e = hipStreamSynchronize(ctx->last_stream);
            if (e != hipSuccess) return ctx->hip_fail("sgm_set_option: synchronize", e);
            const size_t npx = (size_t)ctx->max_w * ctx->max_h;
            const size_t census_bytes = 2 * ctx->census_half * 8 + npx * edge_bytes_per_px(ctx->consts.wide_costs != 0);
            const size_t cvol_bytes = npx * (size_t)dpl_for(ctx->p.num_disparities) * kGroup;
            void *cb = nullptr, *cp = nullptr, *vb = nullptr, *vp = nullptr;
            e = alloc_at_low_word(census_bytes, 0x80000100u, &cb, &cp);
            if (e == hipSuccess && cvol_bytes <= kMaxViewBytes) e = alloc_at_low_word(cvol_bytes, 0x90000100u, &vb, &vp);
            if (e != hipSuccess) {
                dfree(cb);
                dfree(vb);
                return ctx->hip_fail("sgm_set_option: placed allocation", e);
            }
            dfree(ctx->d_census);
            dfree(ctx->d_cvol);
            ctx->d_census_block = cb;
            ctx->d_census = static_cast<uint64_t*>(cp);
            ctx->d_cvol_block = vb;
            ctx->d_cvol = static_cast<uint8_t*>(vp);
            ctx->bytes += 2 * (size_t(1) << 32);
            return SGM_OK;
        }
        case SGM_OPT_COST_VOLUME: {
            const size_t n = (size_t)ctx->max_w * ctx->max_h * (size_t)dpl_for(ctx->p.num_disparities) * kGroup;
            if (value && n > kMaxViewBytes)
                return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "cost volume of " + std::to_string(n) + " bytes exceeds 4 GiB - 1");
            if (value && !ctx->d_cvol) {
                const hipError_t e = dalloc(&ctx->d_cvol, n, ctx->bytes);
                if (e != hipSuccess) return ctx->hip_fail("sgm_set_option: cost volume allocation", e);
            }
            ctx->use_cvol = value != 0;
            return SGM_OK;
        }
        case SGM_OPT_PATH_LINES:
            if (value != 0 && value != 4 && value != 8 && value != 16)
                return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "path lines per block must be 0 (chosen per pair), 4, 8 or 16");
            ctx->path_lines = (int)value;
            return SGM_OK;
        case SGM_OPT_PATH_GROUP:
            if (value != 0 && value != 16 && value != 32 && value != 64)
                return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "lanes per path scanline must be 0 (chosen per pair), 16, 32 or 64");
            ctx->path_group = (int)value;
            return SGM_OK;
        case SGM_OPT_ROW_THREADS:
            if (value != 0 && value != 256 && value != 1024)
                return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "threads per row block must be 0 (chosen per pair), 256 or 1024");
            ctx->row_threads = (int)value;
            return SGM_OK;
        case SGM_OPT_RECOMPUTE_DIRS:
            if (value < 0 || value > 8)
                return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "directions recomputing their costs must be 0..8");
            ctx->consts.cv_from_dir = (int32_t)value;
            return SGM_OK;
        default:
            return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "unknown option " + std::to_string(option));
    }
}

int sgm_dev_malloc(size_t bytes, void** out) {
    if (!out) return SGM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    const hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_dev_malloc", e);
        return SGM_ERR_IO;
    }
    return SGM_OK;
}

int sgm_dev_free(void* ptr) {
    const hipError_t e = hipFree(ptr);
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_dev_free", e);
        return SGM_ERR_IO;
    }
    return SGM_OK;
}

int sgm_dev_upload(void* dst, const void* src, size_t bytes) {
    const hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_dev_upload", e);
        return SGM_ERR_IO;
    }
    return SGM_OK;
}

int sgm_dev_download(void* dst, const void* src, size_t bytes) {
    const hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_dev_download", e);
        return SGM_ERR_IO;
    }
    return SGM_OK;
}

int sgm_dev_memset(void* dst, int value, size_t bytes) {
    const hipError_t e = hipMemset(dst, value, bytes);
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_dev_memset", e);
        return SGM_ERR_IO;
    }
    return SGM_OK;
}

int sgm_dev_copy(void* dst, const void* src, size_t bytes) {
    if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src) | bytes) & 15) {
        g_last_error = "sgm_dev_copy: pointers and size must be multiples of 16 bytes";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    hipError_t e = launch_copy(dst, src, bytes, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_dev_copy", e);
        return SGM_ERR_IO;
    }
    return SGM_OK;
}

int sgm_host_alloc(size_t bytes, void** out) {
    if (!out) return SGM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    const hipError_t e = hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_host_alloc", e);
        return SGM_ERR_IO;
    }
    return SGM_OK;
}

int sgm_host_free(void* ptr) {
    const hipError_t e = hipHostFree(ptr);
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_host_free", e);
        return SGM_ERR_IO;
    }
    return SGM_OK;
}

int sgm_dev_mem_info(size_t* free_bytes, size_t* total_bytes) {
    size_t f = 0, t = 0;
    const hipError_t e = hipMemGetInfo(&f, &t);
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_dev_mem_info", e);
        return SGM_ERR_IO;
    }
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return SGM_OK;
}

// Every call on a context uses the context's scratch buffers, so a call enqueued on another stream
// than the previous call waits for that stream's work first (calls on one stream are ordered anyway).
static hipError_t order_after_last(sgm_context* ctx, hipStream_t s) {
    if (!ctx->last_stream || ctx->last_stream == s) return hipSuccess;
    hipError_t e = hipEventRecord(ctx->order_ev, ctx->last_stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, ctx->order_ev, 0);
    return e;
}

static int check_size(sgm_context* ctx, int width, int height) {
    if (width < 1 || height < 1 || width > ctx->max_w || height > ctx->max_h)
        return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "image size " + std::to_string(width) + "x" + std::to_string(height) +
                                                       " outside the context's 1.." + std::to_string(ctx->max_w) + " x 1.." +
                                                       std::to_string(ctx->max_h));
    return SGM_OK;
}

// Builds the descriptor for one pair over the context's scratch; outputs default to the context's.
static PairDesc context_desc(sgm_context* ctx, const uint8_t* dl, const uint8_t* dr, int64_t pitch, int w, int h,
                             const sgm_outputs* o) {
    PairDesc d{};
    d.left = dl;
    d.right = dr;
    d.pitch = pitch;
    d.census_l = ctx->d_census;
    d.census_r = ctx->d_census + ctx->census_half;
    d.edge = reinterpret_cast<uint32_t*>(ctx->d_census + 2 * ctx->census_half);
    d.lvol = ctx->d_lvol;
    d.cvol = ctx->use_cvol ? ctx->d_cvol : nullptr;
    set_geometry(d, w, h, ctx->p.min_disparity, ctx->p.num_disparities, ctx->ndirs, ctx->consts.wide_costs ? 2 : 1);
    d.disp = (o && o->disparity) ? o->disparity : ctx->d_disp;
    if (o) {
        d.median = o->disparity_median;
        d.conf = o->confidence;
        d.dint = o->disparity_int;
        d.status = o->status;
        d.rdisp = o->right_disparity;
        d.best = o->best_cost;
        d.second = o->second_cost;
        d.sum_volume = o->sum_volume;
        d.cost_volume = o->cost_volume;
    }
    return d;
}

// Adds one completed event set to the running totals; the newest also becomes last_timings.
static void account(sgm_context* ctx, Stages& st, bool newest) {
    float ms[4] = {0, 0, 0, 0}, all = 0, rect = 0, refine = 0, filter = 0, pts = 0;
    for (int i = 0; i < 4; ++i) (void)hipEventElapsedTime(&ms[i], st.ev[i], st.ev[i + 1]);
    if (st.front) (void)hipEventElapsedTime(&rect, st.ev[5], st.ev[0]);
    if (st.post) {
        (void)hipEventElapsedTime(&refine, st.ev[6], st.ev[7]);
        (void)hipEventElapsedTime(&filter, st.ev[7], st.ev[8]);
    }
    if (st.points) (void)hipEventElapsedTime(&pts, st.ev[8], st.ev[9]);
    (void)hipEventElapsedTime(&all, st.front ? st.ev[5] : st.ev[0], st.ev[st.last()]);
    ctx->totals.points_ms += pts;
    ctx->totals.rectify_ms += rect;
    ctx->totals.refine_ms += refine;
    ctx->totals.filter_ms += filter;
    ctx->totals.census_ms += ms[0];
    ctx->totals.paths_ms += ms[1];
    ctx->totals.rows_ms += ms[2];
    ctx->totals.median_ms += ms[3];
    ctx->totals.gpu_ms += all;
    ctx->totals.count += 1;
    if (newest) {
        ctx->timings.census_ms = ms[0];
        ctx->timings.paths_ms = ms[1];
        ctx->timings.rows_ms = ms[2];
        ctx->timings.median_ms = ms[3];
        ctx->timings.gpu_ms = all;
        ctx->timings.total_ms = ctx->pending_wall_ms > 0 ? ctx->pending_wall_ms : all;
        ctx->timings.rectify_ms = rect;
        ctx->timings.refine_ms = refine;
        ctx->timings.filter_ms = filter;
        ctx->timings.points_ms = pts;
    }
    st.recorded = false;
    st.front = false;
    st.post = false;
    st.points = false;
}

// Reads every recorded set (the stream has been synchronised).
static void read_timings(sgm_context* ctx) {
    while (ctx->ring_pending > 0) {
        const int i = (ctx->ring_head - ctx->ring_pending + kEventRing) % kEventRing;
        --ctx->ring_pending;
        account(ctx, ctx->ring[i], ctx->ring_pending == 0);
    }
}

// The event set for the next call; when all are in flight, waits for the oldest and reads it.
static Stages* next_stages(sgm_context* ctx) {
    if (ctx->ring_pending == kEventRing) {
        const int i = (ctx->ring_head - ctx->ring_pending + kEventRing) % kEventRing;
        (void)hipEventSynchronize(ctx->ring[i].ev[ctx->ring[i].last()]);
        --ctx->ring_pending;
        account(ctx, ctx->ring[i], false);
    }
    Stages* st = &ctx->ring[ctx->ring_head];
    ctx->ring_head = (ctx->ring_head + 1) % kEventRing;
    ++ctx->ring_pending;
    return st;
}

int sgm_compute_device(sgm_context* ctx, const uint8_t* d_left, const uint8_t* d_right, ptrdiff_t pitch, int width,
                       int height, const sgm_outputs* d_out, void* stream) {
    if (!ctx) {
        g_last_error = "ctx is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = check_size(ctx, width, height); rc != SGM_OK) return rc;
    if (!d_left || !d_right || pitch < width) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "null image or pitch < width");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    if (const hipError_t oe = order_after_last(ctx, s); oe != hipSuccess) return ctx->hip_fail("stream order", oe);
    PairDesc d = context_desc(ctx, d_left, d_right, pitch, width, height, d_out);
    if (d_out && (d_out->census_left || d_out->census_right))
        return ctx->fail(SGM_ERR_UNSUPPORTED, "census outputs are only available from sgm_compute");
    const bool padded = d.dp != d.nd;
    const hipError_t e = enqueue_pair(d, ctx->consts, padded, s, next_stages(ctx), ctx->path_lines, ctx->path_group, ctx->row_threads);
    if (e != hipSuccess) return ctx->hip_fail("sgm_compute_device: launch", e);
    ctx->last_stream = s;
    ctx->pending_wall_ms = 0;
    return SGM_OK;
}

void* sgm_get_stream(sgm_context* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int sgm_synchronize(sgm_context* ctx) {
    if (!ctx) return SGM_ERR_INVALID_ARGUMENT;
    hipStream_t s = ctx->last_stream ? ctx->last_stream : ctx->stream;
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return ctx->hip_fail("sgm_synchronize", e);
    read_timings(ctx);
    return SGM_OK;
}

int sgm_last_timings(const sgm_context* ctx, sgm_timings* out) {
    if (!ctx || !out) return SGM_ERR_INVALID_ARGUMENT;
    *out = ctx->timings;
    return SGM_OK;
}

int sgm_get_stage_totals(const sgm_context* ctx, sgm_stage_totals* out) {
    if (!ctx || !out) return SGM_ERR_INVALID_ARGUMENT;
    *out = ctx->totals;
    return SGM_OK;
}

int sgm_reset_stage_totals(sgm_context* ctx) {
    if (!ctx) return SGM_ERR_INVALID_ARGUMENT;
    ctx->totals = sgm_stage_totals{};
    return SGM_OK;
}

}  // extern "C"

// Rectifies raw device frames with the resident maps into rect_l / rect_r (pitch map_w / 2).
static hipError_t enqueue_rectify(sgm_context* ctx, const uint8_t* raw_l, const uint8_t* raw_r, int64_t raw_pitch,
                                  uint8_t* rect_l, uint8_t* rect_r, hipStream_t s, Stages* st) {
    RectifyDesc r{};
    const size_t n = (size_t)ctx->map_w * ctx->map_h;
    r.raw[0] = raw_l;
    r.raw[1] = raw_r;
    r.map[0] = ctx->d_maps;
    r.map[1] = ctx->d_maps + 2 * n;
    r.out[0] = rect_l;
    r.out[1] = rect_r;
    r.raw_pitch = raw_pitch;
    r.raw_w = ctx->raw_w;
    r.raw_h = ctx->raw_h;
    r.map_w = ctx->map_w;
    r.out_w = ctx->map_w / 2;
    r.out_h = ctx->map_h / 2;
    r.out_pitch = r.out_w;
    hipError_t e;
    if (st && !st->front) {
        if ((e = hipEventRecord(st->ev[5], s)) != hipSuccess) return e;
        st->front = true;
    }
    return launch_rectify(r, s);
}

static int check_raw(sgm_context* ctx, int raw_w, int raw_h) {
    if (!ctx->d_maps) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "set_rectification() not called");
    if (raw_w != ctx->raw_w || raw_h != ctx->raw_h) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "raw size mismatch");
    return SGM_OK;
}

// Host frames in, host outputs back, blocking: the shared body of sgm_compute (rectified input,
// raw == false) and sgm_compute_raw (raw frames rectified on the device first).
struct HostFrame {
    const uint8_t* left;
    const uint8_t* right;
    ptrdiff_t stride;
    bool raw;
    uint8_t* rect_left;   // raw only: rectified pair back to the host, may be null
    uint8_t* rect_right;
};

static int compute_host(sgm_context* ctx, const HostFrame& f, int width, int height, const sgm_outputs* out,
                        const char* who) {
    const auto t0 = std::chrono::steady_clock::now();
    const std::string w(who);
    if (!out || !out->disparity) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "out->disparity is required");
    const size_t npx = (size_t)width * height;
    const size_t nvol = npx * (size_t)ctx->p.num_disparities;
    hipStream_t s = ctx->stream;
    if (const hipError_t oe = order_after_last(ctx, s); oe != hipSuccess) return ctx->hip_fail("stream order", oe);
    hipError_t e = hipSuccess;
    if (out->sum_volume && !ctx->d_sum) e = dalloc(&ctx->d_sum, (size_t)ctx->max_w * ctx->max_h * ctx->p.num_disparities, ctx->bytes);
    if (e == hipSuccess && out->cost_volume && !ctx->d_cost)
        e = dalloc(&ctx->d_cost, (size_t)ctx->max_w * ctx->max_h * ctx->p.num_disparities, ctx->bytes);
    if (e != hipSuccess) return ctx->hip_fail((w + ": debug volume allocation").c_str(), e);
    uint8_t* dl = ctx->d_img;
    uint8_t* dr = ctx->d_img + (size_t)ctx->max_w * ctx->max_h;
    Stages* st = next_stages(ctx);
    if (f.raw) {
        const size_t rpx = (size_t)ctx->raw_w * ctx->raw_h;
        e = hipMemcpy2DAsync(ctx->d_raw, ctx->raw_w, f.left, f.stride, ctx->raw_w, ctx->raw_h, hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(ctx->d_raw + rpx, ctx->raw_w, f.right, f.stride, ctx->raw_w, ctx->raw_h, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return ctx->hip_fail((w + ": upload").c_str(), e);
        if ((e = enqueue_rectify(ctx, ctx->d_raw, ctx->d_raw + rpx, ctx->raw_w, dl, dr, s, st)) != hipSuccess)
            return ctx->hip_fail((w + ": rectify").c_str(), e);
    } else {
        e = hipMemcpy2DAsync(dl, width, f.left, f.stride, width, height, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpy2DAsync(dr, width, f.right, f.stride, width, height, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return ctx->hip_fail((w + ": upload").c_str(), e);
    }
    sgm_outputs o{};
    o.disparity = ctx->d_disp;
    o.disparity_median = out->disparity_median ? ctx->d_median : nullptr;
    o.confidence = out->confidence ? ctx->d_conf : nullptr;
    o.disparity_int = out->disparity_int ? ctx->d_dint : nullptr;
    o.status = out->status ? ctx->d_status : nullptr;
    o.right_disparity = out->right_disparity ? ctx->d_rdisp : nullptr;
    o.best_cost = out->best_cost ? ctx->d_best : nullptr;
    o.second_cost = out->second_cost ? ctx->d_second : nullptr;
    o.sum_volume = out->sum_volume ? ctx->d_sum : nullptr;
    o.cost_volume = out->cost_volume ? ctx->d_cost : nullptr;
    PairDesc d = context_desc(ctx, dl, dr, width, width, height, &o);
    e = enqueue_pair(d, ctx->consts, d.dp != d.nd, s, st, ctx->path_lines, ctx->path_group, ctx->row_threads);
    if (e != hipSuccess) return ctx->hip_fail((w + ": launch").c_str(), e);
    struct Copy {
        void* host;
        const void* dev;
        size_t bytes;
    };
    const Copy copies[] = {
        {out->disparity, ctx->d_disp, npx * 4},
        {out->disparity_median, ctx->d_median, npx * 4},
        {out->confidence, ctx->d_conf, npx * 4},
        {out->disparity_int, ctx->d_dint, npx * 4},
        {out->status, ctx->d_status, npx},
        {out->right_disparity, ctx->d_rdisp, npx * 4},
        {out->best_cost, ctx->d_best, npx * 2},
        {out->second_cost, ctx->d_second, npx * 2},
        {out->sum_volume, ctx->d_sum, nvol * 2},
        {out->cost_volume, ctx->d_cost, nvol * 2},
        {f.raw ? f.rect_left : nullptr, dl, npx},
        {f.raw ? f.rect_right : nullptr, dr, npx},
    };
    for (const Copy& c : copies)
        if (c.host && (e = hipMemcpyAsync(c.host, c.dev, c.bytes, hipMemcpyDeviceToHost, s)) != hipSuccess)
            return ctx->hip_fail((w + ": download").c_str(), e);
    for (const auto& [host, dev] : {std::pair<uint64_t*, const uint64_t*>{out->census_left, d.census_l},
                                    std::pair<uint64_t*, const uint64_t*>{out->census_right, d.census_r}})
        if (host && (e = hipMemcpy2DAsync(host, (size_t)width * 8, dev + d.census_padl, (size_t)d.census_pitch * 8,
                                          (size_t)width * 8, height, hipMemcpyDeviceToHost, s)) != hipSuccess)
            return ctx->hip_fail((w + ": census download").c_str(), e);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return ctx->hip_fail((w + ": execution").c_str(), e);
    ctx->last_stream = s;
    ctx->pending_wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    read_timings(ctx);
    return SGM_OK;
}

extern "C" {

int sgm_compute(sgm_context* ctx, const uint8_t* left, const uint8_t* right, ptrdiff_t stride, int width, int height,
                const sgm_outputs* out) {
    if (!ctx) {
        g_last_error = "ctx is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = check_size(ctx, width, height); rc != SGM_OK) return rc;
    if (!left || !right || stride < width) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "null image or stride < width");
    return compute_host(ctx, HostFrame{left, right, stride, false, nullptr, nullptr}, width, height, out, "sgm_compute");
}

int sgm_set_rectification(sgm_context* ctx, const float* left_map_x, const float* left_map_y, const float* right_map_x,
                          const float* right_map_y, int map_w, int map_h, int raw_w, int raw_h) {
    if (!ctx) {
        g_last_error = "ctx is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (!left_map_x || !left_map_y || !right_map_x || !right_map_y) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "null map");
    if (map_w < 2 || map_h < 2 || map_w / 2 > ctx->stereo_w || map_h / 2 > ctx->stereo_h)
        return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "remap resolution must be twice the stereo resolution (stereo size " +
                                                       std::to_string(map_w / 2) + "x" + std::to_string(map_h / 2) +
                                                       " outside the context's " + std::to_string(ctx->stereo_w) + "x" +
                                                       std::to_string(ctx->stereo_h) + ")");
    // Raw sizes stay below 2^24 so that the kernel's float bounds test is exact.
    if (raw_w < 1 || raw_h < 1 || raw_w >= (1 << 24) || raw_h >= (1 << 24))
        return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "raw size must be in [1, 2^24)");
    // Work still reading the old maps may be on the context's stream or on the last caller stream.
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && ctx->last_stream) e = hipStreamSynchronize(ctx->last_stream);
    if (e != hipSuccess) return ctx->hip_fail("sgm_set_rectification: synchronize", e);
    dfree(ctx->d_maps);
    dfree(ctx->d_raw);
    ctx->bytes -= ctx->maps_bytes + ctx->raw_bytes;
    ctx->d_maps = nullptr;
    ctx->d_raw = nullptr;
    ctx->maps_bytes = ctx->raw_bytes = 0;
    ctx->map_w = ctx->map_h = ctx->raw_w = ctx->raw_h = 0;
    const size_t n = (size_t)map_w * map_h;
    const size_t rpx = (size_t)raw_w * raw_h;
    size_t got = 0;
    if ((e = dalloc(&ctx->d_maps, 4 * n, got)) != hipSuccess) return ctx->hip_fail("sgm_set_rectification: map allocation", e);
    ctx->maps_bytes = got;
    got = 0;
    if ((e = dalloc(&ctx->d_raw, 2 * rpx, got)) != hipSuccess) {
        dfree(ctx->d_maps);
        ctx->d_maps = nullptr;
        ctx->maps_bytes = 0;
        return ctx->hip_fail("sgm_set_rectification: raw frame allocation", e);
    }
    ctx->raw_bytes = got;
    ctx->bytes += ctx->maps_bytes + ctx->raw_bytes;
    // The kernel reads the (x, y) pairs of two neighbouring map pixels as one 16-byte load.
    std::vector<float> packed(2 * n);
    const float* mx[2] = {left_map_x, right_map_x};
    const float* my[2] = {left_map_y, right_map_y};
    for (int side = 0; side < 2 && e == hipSuccess; ++side) {
        for (size_t i = 0; i < n; ++i) {
            packed[2 * i] = mx[side][i];
            packed[2 * i + 1] = my[side][i];
        }
        e = hipMemcpy(ctx->d_maps + 2 * n * side, packed.data(), 8 * n, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        dfree(ctx->d_maps);
        dfree(ctx->d_raw);
        ctx->bytes -= ctx->maps_bytes + ctx->raw_bytes;
        ctx->d_maps = nullptr;
        ctx->d_raw = nullptr;
        ctx->maps_bytes = ctx->raw_bytes = 0;
        return ctx->hip_fail("sgm_set_rectification: map upload", e);
    }
    ctx->map_w = map_w;
    ctx->map_h = map_h;
    ctx->raw_w = raw_w;
    ctx->raw_h = raw_h;
    return SGM_OK;
}

int sgm_compute_raw(sgm_context* ctx, const uint8_t* raw_left, const uint8_t* raw_right, ptrdiff_t raw_stride, int raw_w,
                    int raw_h, const sgm_outputs* out, uint8_t* rect_left, uint8_t* rect_right) {
    if (!ctx) {
        g_last_error = "ctx is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = check_raw(ctx, raw_w, raw_h); rc != SGM_OK) return rc;
    if (!raw_left || !raw_right || raw_stride < raw_w) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "null image or stride < width");
    if (int rc = check_size(ctx, ctx->map_w / 2, ctx->map_h / 2); rc != SGM_OK) return rc;
    return compute_host(ctx, HostFrame{raw_left, raw_right, raw_stride, true, rect_left, rect_right}, ctx->map_w / 2,
                        ctx->map_h / 2, out, "sgm_compute_raw");
}

int sgm_compute_raw_device(sgm_context* ctx, const uint8_t* d_raw_left, const uint8_t* d_raw_right, ptrdiff_t raw_pitch,
                           int raw_w, int raw_h, const sgm_outputs* d_out, uint8_t* d_rect_left, uint8_t* d_rect_right,
                           void* stream) {
    if (!ctx) {
        g_last_error = "ctx is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = check_raw(ctx, raw_w, raw_h); rc != SGM_OK) return rc;
    if (!d_raw_left || !d_raw_right || raw_pitch < raw_w) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "null image or pitch < width");
    if (d_out && (d_out->census_left || d_out->census_right))
        return ctx->fail(SGM_ERR_UNSUPPORTED, "census outputs are only available from sgm_compute");
    const int width = ctx->map_w / 2, height = ctx->map_h / 2;
    if (int rc = check_size(ctx, width, height); rc != SGM_OK) return rc;
    uint8_t* dl = d_rect_left ? d_rect_left : ctx->d_img;
    uint8_t* dr = d_rect_right ? d_rect_right : ctx->d_img + (size_t)ctx->max_w * ctx->max_h;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    if (const hipError_t oe = order_after_last(ctx, s); oe != hipSuccess) return ctx->hip_fail("stream order", oe);
    Stages* st = next_stages(ctx);
    hipError_t e = enqueue_rectify(ctx, d_raw_left, d_raw_right, raw_pitch, dl, dr, s, st);
    if (e != hipSuccess) return ctx->hip_fail("sgm_compute_raw_device: rectify", e);
    PairDesc d = context_desc(ctx, dl, dr, width, width, height, d_out);
    if ((e = enqueue_pair(d, ctx->consts, d.dp != d.nd, s, st, ctx->path_lines, ctx->path_group, ctx->row_threads)) != hipSuccess)
        return ctx->hip_fail("sgm_compute_raw_device: launch", e);
    ctx->last_stream = s;
    ctx->pending_wall_ms = 0;
    return SGM_OK;
}

}  // extern "C"

// ----------------------------------------------------------------------------- compute_disparity
// stereo.cpp:496-536: pyramid, SGM at the top level, refine_level down to level 0, refine_detail on
// the full-resolution pair, filter_disparity, remove_speckles, confidence. fl / fr are null or the
// 2w x 2h full-resolution pair.

// out_disp / out_conf null: the results stay in the context's buffers, returned in *res_disp /
// *res_conf (the finalize kernel reads and writes each pixel in one thread, so it may work in place).
static hipError_t enqueue_disparity(sgm_context* ctx, const uint8_t* dl, const uint8_t* dr, int64_t pitch, int w, int h,
                                    const uint8_t* fl, const uint8_t* fr, int64_t fpitch, float* out_disp,
                                    float* out_conf, hipStream_t s, Stages* st, float** res_disp = nullptr,
                                    float** res_conf = nullptr) {
    const sgm_stereo_params& sp = ctx->sp;
    const int L = ctx->levels;
    int lw[kMaxLevels + 1], lh[kMaxLevels + 1];
    const uint8_t* il[kMaxLevels + 1];
    const uint8_t* ir[kMaxLevels + 1];
    int64_t ip[kMaxLevels + 1];
    lw[0] = w;
    lh[0] = h;
    il[0] = dl;
    ir[0] = dr;
    ip[0] = pitch;
    hipError_t e;
    if (st && L > 0 && !st->front) {
        if ((e = hipEventRecord(st->ev[5], s)) != hipSuccess) return e;
        st->front = true;
    }
    size_t off = 0;
    for (int i = 1; i <= L; ++i) {
        lw[i] = lw[i - 1] / 2;
        lh[i] = lh[i - 1] / 2;
        uint8_t* a = ctx->d_pyr + off;
        uint8_t* b = a + (size_t)lw[i] * lh[i];
        off += 2 * (size_t)lw[i] * lh[i];
        if ((e = launch_downsample2(il[i - 1], ir[i - 1], ip[i - 1], lw[i - 1], lh[i - 1], a, b, lw[i], s)) != hipSuccess)
            return e;
        il[i] = a;
        ir[i] = b;
        ip[i] = lw[i];
    }
    sgm_outputs o{};
    o.disparity = ctx->d_disp;
    o.disparity_median = L > 0 ? ctx->d_wm : nullptr;  // refine_level's median3 of the coarse disparity
    PairDesc d = context_desc(ctx, il[L], ir[L], ip[L], lw[L], lh[L], &o);
    if ((e = enqueue_pair(d, ctx->consts, d.dp != d.nd, s, st, ctx->path_lines, ctx->path_group, ctx->row_threads)) != hipSuccess) return e;
    if (st) {
        if ((e = hipEventRecord(st->ev[6], s)) != hipSuccess) return e;
        st->post = true;
    }
    float* cur = ctx->d_disp;
    const float* conf_src = nullptr;
    for (int l = L - 1; l >= 0; --l) {
        RefineDesc R{};
        R.left = il[l];
        R.right = ir[l];
        R.pitch = ip[l];
        R.w = lw[l];
        R.h = lh[l];
        R.coarse = ctx->d_wm;
        R.cw = lw[l + 1];
        R.ch = lh[l + 1];
        R.disp = ctx->d_wa;
        R.conf = ctx->d_wc;
        R.radius = sp.refine_zncc_radius;
        R.search_radius = sp.refine_search_radius;
        R.reject_edge = sp.refine_reject_edge_winners ? 1 : 0;
        R.subpixel = ctx->p.subpixel ? 1 : 0;
        R.min_zncc = sp.refine_min_zncc;
        if ((e = launch_refine(R, s)) != hipSuccess) return e;
        if (l > 0 && (e = launch_median3(ctx->d_wa, ctx->d_wm, lw[l], lh[l], s)) != hipSuccess) return e;
        cur = ctx->d_wa;
        conf_src = ctx->d_wc;
    }
    // Ping-pong partner of the current map (d_disp, d_wa and d_wb hold disparity maps).
    auto other = [&](const float* c) { return c == ctx->d_wa ? ctx->d_wb : ctx->d_wa; };
    if (sp.detail_enabled && fl && fr) {
        DetailDesc D{};
        D.fl = fl;
        D.fr = fr;
        D.fpitch = fpitch;
        D.W = 2 * w;
        D.H = 2 * h;
        D.first = cur;
        D.out = other(cur);
        D.w = w;
        D.h = h;
        D.radius = sp.detail_radius;
        D.K = std::min(16, (int)std::lround(sp.detail_search / sp.detail_step));
        D.subpixel = ctx->p.subpixel ? 1 : 0;
        D.step = sp.detail_step;
        D.min_zncc = sp.detail_min_zncc;
        if ((e = launch_detail(D, s)) != hipSuccess) return e;
        cur = D.out;
    }
    if (st && (e = hipEventRecord(st->ev[7], s)) != hipSuccess) return e;
    if (sp.filter_occlusion && (e = launch_occlusion(cur, w, h, sp.filter_occlusion_tolerance, s)) != hipSuccess) return e;
    if (sp.filter_spikes) {
        float* next = other(cur);
        if ((e = launch_spikes(cur, next, w, h, sp.filter_spike_threshold, s)) != hipSuccess) return e;
        cur = next;
    }
    const bool speckle = sp.speckle_max_region_size > 1;
    if (speckle && (e = launch_ccl(cur, ctx->d_labels, ctx->d_local, ctx->d_sizes, w, h, sp.speckle_max_diff, s)) != hipSuccess) return e;
    if (!out_disp) out_disp = cur;
    if (!out_conf && res_conf) out_conf = conf_src ? ctx->d_wc : (ctx->d_wc ? ctx->d_wc : ctx->d_conf);
    if ((e = launch_finalize(cur, speckle ? ctx->d_labels : nullptr, ctx->d_sizes, sp.speckle_max_region_size, conf_src,
                             out_disp, out_conf, w, h, s)) != hipSuccess)
        return e;
    if (res_disp) *res_disp = out_disp;
    if (res_conf) *res_conf = out_conf;
    if (st && (e = hipEventRecord(st->ev[8], s)) != hipSuccess) return e;
    if (st) st->recorded = true;
    return hipSuccess;
}

static int check_stereo_size(sgm_context* ctx, int width, int height) {
    if (width < 1 || height < 1 || width > ctx->stereo_w || height > ctx->stereo_h || (width >> ctx->levels) < 1 ||
        (height >> ctx->levels) < 1)
        return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "image size " + std::to_string(width) + "x" + std::to_string(height) +
                                                       " outside the context's " + std::to_string(ctx->stereo_w) + "x" +
                                                       std::to_string(ctx->stereo_h));
    return SGM_OK;
}

// Raw frames (device) through the resident maps into the stereo-resolution pair rl / rr (pitch w),
// and, when the detail refinement runs, into the full-resolution pair in d_full first.
static hipError_t enqueue_raw_front(sgm_context* ctx, const uint8_t* raw_l, const uint8_t* raw_r, int64_t raw_pitch,
                                    uint8_t* rl, uint8_t* rr, hipStream_t s, Stages* st, const uint8_t** fl,
                                    const uint8_t** fr) {
    const int w = ctx->map_w / 2, h = ctx->map_h / 2;
    *fl = *fr = nullptr;
    if (!(ctx->sp.detail_enabled && ctx->d_full)) return enqueue_rectify(ctx, raw_l, raw_r, raw_pitch, rl, rr, s, st);
    hipError_t e;
    if (st && !st->front) {
        if ((e = hipEventRecord(st->ev[5], s)) != hipSuccess) return e;
        st->front = true;
    }
    RectifyDesc r{};
    const size_t n = (size_t)ctx->map_w * ctx->map_h;
    r.raw[0] = raw_l;
    r.raw[1] = raw_r;
    r.map[0] = ctx->d_maps;
    r.map[1] = ctx->d_maps + 2 * n;
    r.raw_pitch = raw_pitch;
    r.raw_w = ctx->raw_w;
    r.raw_h = ctx->raw_h;
    r.map_w = ctx->map_w;
    uint8_t* f0 = ctx->d_full;
    uint8_t* f1 = ctx->d_full + (size_t)4 * w * h;
    if ((e = launch_rectify_full(r, f0, f1, 2 * w, 2 * w, 2 * h, s)) != hipSuccess) return e;
    if ((e = launch_downsample2(f0, f1, 2 * w, 2 * w, 2 * h, rl, rr, w, s)) != hipSuccess) return e;
    *fl = f0;
    *fr = f1;
    return hipSuccess;
}

// Host frames in, host maps out, blocking: sgm_compute_disparity (raw == false) and
// sgm_compute_disparity_raw (raw == true: left / right are raw frames).
static int disparity_host(sgm_context* ctx, const uint8_t* left, const uint8_t* right, ptrdiff_t stride, int w, int h,
                          const uint8_t* full_left, const uint8_t* full_right, ptrdiff_t full_stride, bool full, bool raw,
                          float* disparity, float* confidence, uint8_t* rect_l, uint8_t* rect_r, const char* who) {
    const auto t0 = std::chrono::steady_clock::now();
    const std::string nm(who);
    hipStream_t s = ctx->stream;
    if (const hipError_t oe = order_after_last(ctx, s); oe != hipSuccess) return ctx->hip_fail("stream order", oe);
    const size_t npx = (size_t)w * h;
    uint8_t* dl = ctx->d_img;
    uint8_t* dr = ctx->d_img + npx;
    const uint8_t* fl = nullptr;
    const uint8_t* fr = nullptr;
    Stages* st = next_stages(ctx);
    hipError_t e = hipSuccess;
    if (raw) {
        const size_t rpx = (size_t)ctx->raw_w * ctx->raw_h;
        e = hipMemcpy2DAsync(ctx->d_raw, ctx->raw_w, left, stride, ctx->raw_w, ctx->raw_h, hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(ctx->d_raw + rpx, ctx->raw_w, right, stride, ctx->raw_w, ctx->raw_h, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return ctx->hip_fail((nm + ": upload").c_str(), e);
        if ((e = enqueue_raw_front(ctx, ctx->d_raw, ctx->d_raw + rpx, ctx->raw_w, dl, dr, s, st, &fl, &fr)) != hipSuccess)
            return ctx->hip_fail((nm + ": rectify").c_str(), e);
    } else {
        e = hipMemcpy2DAsync(dl, w, left, stride, w, h, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpy2DAsync(dr, w, right, stride, w, h, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && full) {
            uint8_t* f0 = ctx->d_full;
            uint8_t* f1 = ctx->d_full + 4 * npx;
            e = hipMemcpy2DAsync(f0, 2 * w, full_left, full_stride, 2 * w, 2 * h, hipMemcpyHostToDevice, s);
            if (e == hipSuccess) e = hipMemcpy2DAsync(f1, 2 * w, full_right, full_stride, 2 * w, 2 * h, hipMemcpyHostToDevice, s);
            fl = f0;
            fr = f1;
        }
        if (e != hipSuccess) return ctx->hip_fail((nm + ": upload").c_str(), e);
    }
    float* dd = nullptr;
    float* dc = nullptr;
    if ((e = enqueue_disparity(ctx, dl, dr, w, w, h, fl, fr, 2 * w, nullptr, nullptr, s, st, &dd, &dc)) != hipSuccess)
        return ctx->hip_fail((nm + ": launch").c_str(), e);
    if ((e = hipMemcpyAsync(disparity, dd, npx * 4, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return ctx->hip_fail((nm + ": download").c_str(), e);
    if (confidence && (e = hipMemcpyAsync(confidence, dc, npx * 4, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return ctx->hip_fail((nm + ": download").c_str(), e);
    if (rect_l && (e = hipMemcpyAsync(rect_l, dl, npx, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return ctx->hip_fail((nm + ": download").c_str(), e);
    if (rect_r && (e = hipMemcpyAsync(rect_r, dr, npx, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return ctx->hip_fail((nm + ": download").c_str(), e);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return ctx->hip_fail((nm + ": execution").c_str(), e);
    ctx->last_stream = s;
    ctx->pending_wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    read_timings(ctx);
    return SGM_OK;
}

extern "C" {

int sgm_compute_disparity(sgm_context* ctx, const uint8_t* left, const uint8_t* right, ptrdiff_t stride, int width,
                          int height, const uint8_t* full_left, const uint8_t* full_right, ptrdiff_t full_stride,
                          int full_width, int full_height, float* disparity, float* confidence) {
    if (!ctx) {
        g_last_error = "ctx is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = check_stereo_size(ctx, width, height); rc != SGM_OK) return rc;
    if (!left || !right || stride < width || !disparity)
        return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "null image or output, or stride < width");
    // The detail refinement needs the images left and right were downsampled from (stereo.cpp:524).
    const bool full = ctx->sp.detail_enabled && ctx->d_full && full_left && full_right && full_width == 2 * width &&
                      full_height == 2 * height;
    if (full && full_stride < full_width) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "full_stride < full_width");
    return disparity_host(ctx, left, right, stride, width, height, full_left, full_right, full_stride, full, false,
                          disparity, confidence, nullptr, nullptr, "sgm_compute_disparity");
}

int sgm_compute_disparity_device(sgm_context* ctx, const uint8_t* d_left, const uint8_t* d_right, ptrdiff_t pitch,
                                 int width, int height, const uint8_t* d_full_left, const uint8_t* d_full_right,
                                 ptrdiff_t full_pitch, int full_width, int full_height, float* d_disparity,
                                 float* d_confidence, void* stream) {
    if (!ctx) {
        g_last_error = "ctx is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = check_stereo_size(ctx, width, height); rc != SGM_OK) return rc;
    if (!d_left || !d_right || pitch < width || !d_disparity)
        return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "null image or output, or pitch < width");
    const bool full = ctx->sp.detail_enabled && d_full_left && d_full_right && full_width == 2 * width &&
                      full_height == 2 * height;
    if (full && full_pitch < full_width) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "full_pitch < full_width");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    if (const hipError_t oe = order_after_last(ctx, s); oe != hipSuccess) return ctx->hip_fail("stream order", oe);
    const hipError_t e = enqueue_disparity(ctx, d_left, d_right, pitch, width, height, full ? d_full_left : nullptr,
                                           full ? d_full_right : nullptr, full_pitch, d_disparity, d_confidence, s,
                                           next_stages(ctx));
    if (e != hipSuccess) return ctx->hip_fail("sgm_compute_disparity_device: launch", e);
    ctx->last_stream = s;
    ctx->pending_wall_ms = 0;
    return SGM_OK;
}

int sgm_compute_disparity_raw(sgm_context* ctx, const uint8_t* raw_left, const uint8_t* raw_right, ptrdiff_t raw_stride,
                              int raw_w, int raw_h, float* disparity, float* confidence, uint8_t* rect_left,
                              uint8_t* rect_right) {
    if (!ctx) {
        g_last_error = "ctx is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = check_raw(ctx, raw_w, raw_h); rc != SGM_OK) return rc;
    if (!raw_left || !raw_right || raw_stride < raw_w || !disparity)
        return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "null image or output, or stride < width");
    const int w = ctx->map_w / 2, h = ctx->map_h / 2;
    if (int rc = check_stereo_size(ctx, w, h); rc != SGM_OK) return rc;
    return disparity_host(ctx, raw_left, raw_right, raw_stride, w, h, nullptr, nullptr, 0, false, true, disparity,
                          confidence, rect_left, rect_right, "sgm_compute_disparity_raw");
}

int sgm_compute_disparity_raw_device(sgm_context* ctx, const uint8_t* d_raw_left, const uint8_t* d_raw_right,
                                     ptrdiff_t raw_pitch, int raw_w, int raw_h, float* d_disparity, float* d_confidence,
                                     uint8_t* d_rect_left, uint8_t* d_rect_right, void* stream) {
    if (!ctx) {
        g_last_error = "ctx is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = check_raw(ctx, raw_w, raw_h); rc != SGM_OK) return rc;
    if (!d_raw_left || !d_raw_right || raw_pitch < raw_w || !d_disparity)
        return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "null image or output, or pitch < width");
    const int w = ctx->map_w / 2, h = ctx->map_h / 2;
    if (int rc = check_stereo_size(ctx, w, h); rc != SGM_OK) return rc;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    if (const hipError_t oe = order_after_last(ctx, s); oe != hipSuccess) return ctx->hip_fail("stream order", oe);
    uint8_t* rl = d_rect_left ? d_rect_left : ctx->d_img;
    uint8_t* rr = d_rect_right ? d_rect_right : ctx->d_img + (size_t)w * h;
    Stages* st = next_stages(ctx);
    const uint8_t* fl = nullptr;
    const uint8_t* fr = nullptr;
    hipError_t e = enqueue_raw_front(ctx, d_raw_left, d_raw_right, raw_pitch, rl, rr, s, st, &fl, &fr);
    if (e == hipSuccess)
        e = enqueue_disparity(ctx, rl, rr, w, w, h, fl, fr, 2 * w, d_disparity, d_confidence, s, st);
    if (e != hipSuccess) return ctx->hip_fail("sgm_compute_disparity_raw_device: launch", e);
    ctx->last_stream = s;
    ctx->pending_wall_ms = 0;
    return SGM_OK;
}

}  // extern "C"

// ----------------------------------------------------------------------------- point frames

static PointsDesc points_desc(const sgm_rectified_geometry& g, const sgm_point_params& pp, const float* disp,
                              const float* conf, int w, int h, const sgm_frame& out) {
    PointsDesc P{};
    P.disp = disp;
    P.conf = conf;
    P.points = out.points_xyzw;
    P.normals = out.normals_xyzw;
    P.weights = out.weights;
    P.w = w;
    P.h = h;
    P.f = g.f;
    P.cx = g.cx;
    P.cy = g.cy;
    P.baseline = g.baseline;
    P.cx_offset = g.cx_offset;
    P.min_depth = pp.min_depth;
    P.max_depth = pp.max_depth;
    P.max_depth_jump = pp.max_depth_jump;
    P.edge_weight = pp.edge_weight;
    P.weight_reference_depth = pp.weight_reference_depth;
    P.rc_x = (float)g.baseline;
    return P;
}

static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Raw frames (device) -> rectified pair -> compute_disparity -> point frame in `out` (device).
static hipError_t enqueue_frame(sgm_context* ctx, const uint8_t* raw_l, const uint8_t* raw_r, int64_t raw_pitch,
                                uint8_t* rl, uint8_t* rr, const sgm_frame& out, hipStream_t s, Stages* st) {
    const int w = ctx->map_w / 2, h = ctx->map_h / 2;
    const uint8_t* fl = nullptr;
    const uint8_t* fr = nullptr;
    hipError_t e = enqueue_raw_front(ctx, raw_l, raw_r, raw_pitch, rl, rr, s, st, &fl, &fr);
    float* dd = nullptr;
    float* dc = nullptr;
    if (e == hipSuccess) e = enqueue_disparity(ctx, rl, rr, w, w, h, fl, fr, 2 * w, nullptr, nullptr, s, st, &dd, &dc);
    if (e == hipSuccess) e = launch_points(points_desc(ctx->geom, ctx->pparams, dd, dc, w, h, out), s);
    if (e == hipSuccess && st) {
        e = hipEventRecord(st->ev[9], s);
        st->points = true;
    }
    return e;
}

static int check_frame(sgm_context* ctx) {
    if (!ctx->have_points) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "set_point_params() not called");
    return SGM_OK;
}

extern "C" {

void sgm_point_params_default(sgm_point_params* p) {
    if (!p) return;
    p->min_depth = 150.0f;
    p->max_depth = 700.0f;
    p->max_depth_jump = 4.0f;
    p->edge_weight = 0.25f;
    p->weight_reference_depth = 0.0f;
}

int sgm_disparity_to_points_device(const float* d_disparity, const float* d_confidence, int width, int height,
                                   const sgm_rectified_geometry* geometry, const sgm_point_params* params,
                                   const sgm_frame* d_out, void* stream) {
    if (!d_disparity || !geometry || !params || !d_out || !d_out->points_xyzw || !d_out->normals_xyzw || !d_out->weights ||
        width < 1 || height < 1 || !aligned16(d_out->points_xyzw) || !aligned16(d_out->normals_xyzw)) {
        g_last_error = "sgm_disparity_to_points_device: null pointer, size < 1 or float4 arrays not 16-byte aligned";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    const hipError_t e = launch_points(points_desc(*geometry, *params, d_disparity, d_confidence, width, height, *d_out),
                                       static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_disparity_to_points_device", e);
        return SGM_ERR_IO;
    }
    return SGM_OK;
}

int sgm_pack_depth_device(const float* d_points_xyzw, const float* d_weights, int width, int height, uint8_t* d_out,
                          void* stream) {
    if (!d_points_xyzw || !d_out || width < 1 || height < 1 || !aligned16(d_points_xyzw) ||
        (reinterpret_cast<uintptr_t>(d_out) & 1u)) {
        g_last_error = "sgm_pack_depth_device: null pointer, size < 1 or misaligned buffers";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    const hipError_t e = launch_pack(d_points_xyzw, d_weights, d_out, width, height, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_pack_depth_device", e);
        return SGM_ERR_IO;
    }
    return SGM_OK;
}

size_t sgm_speckle_workspace_bytes(int width, int height) {
    if (width < 1 || height < 1) return 0;
    return (size_t)width * (size_t)height * 12;
}

int sgm_remove_speckles_device(const float* d_disparity, float* d_out, int width, int height, int max_region_size,
                               float max_diff, void* d_work, void* stream) {
    const bool speckle = max_region_size > 1;
    if (!d_disparity || !d_out || width < 1 || height < 1 || (int64_t)width * height > INT32_MAX ||
        (speckle && (!d_work || (reinterpret_cast<uintptr_t>(d_work) & 3u)))) {
        g_last_error = "sgm_remove_speckles_device: null pointer, size < 1, more than 2^31 - 1 pixels, or a missing "
                       "or misaligned workspace";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t npx = (size_t)width * height;
    int32_t* labels = static_cast<int32_t*>(d_work);
    uint32_t* local = reinterpret_cast<uint32_t*>(labels + npx);
    uint32_t* sizes = local + npx;
    hipError_t e = speckle ? launch_ccl(d_disparity, labels, local, sizes, width, height, max_diff, s) : hipSuccess;
    if (e == hipSuccess)
        e = launch_finalize(d_disparity, speckle ? labels : nullptr, sizes, max_region_size, nullptr, d_out, nullptr, width,
                            height, s);
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_remove_speckles_device", e);
        return SGM_ERR_IO;
    }
    return SGM_OK;
}

int sgm_filter_disparity_device(float* d_disparity, int width, int height, int occlusion, float occlusion_tolerance,
                                int spikes, float spike_threshold, float* d_work, void* stream) {
    std::string err;
    if (!d_disparity || width < 1 || height < 1 || width > kMaxWidth || (spikes && !d_work) ||
        (reinterpret_cast<uintptr_t>(d_disparity) & 3u) || (reinterpret_cast<uintptr_t>(d_work) & 3u)) {
        g_last_error = "sgm_filter_disparity_device: null pointer, size < 1, width above " + std::to_string(kMaxWidth) +
                       ", or a missing or misaligned workspace";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = occlusion ? check_row_lds(width, err) : SGM_OK; rc != SGM_OK) {
        g_last_error = "sgm_filter_disparity_device: " + err;
        return rc;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = occlusion ? launch_occlusion(d_disparity, width, height, occlusion_tolerance, s) : hipSuccess;
    if (e == hipSuccess && spikes) {
        e = launch_spikes(d_disparity, d_work, width, height, spike_threshold, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_disparity, d_work, (size_t)width * height * sizeof(float), hipMemcpyDeviceToDevice, s);
    }
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_filter_disparity_device", e);
        return SGM_ERR_IO;
    }
    return SGM_OK;
}

int sgm_set_point_params(sgm_context* ctx, const sgm_rectified_geometry* geometry, const sgm_point_params* params) {
    if (!ctx) {
        g_last_error = "ctx is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (!geometry || !params || !(geometry->f > 0)) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "null parameters or f <= 0");
    if (!ctx->d_pts) {
        const size_t spx = (size_t)ctx->stereo_w * ctx->stereo_h;
        size_t got = 0;
        hipError_t e = dalloc(&ctx->d_pts, 4 * spx, got);
        if (e == hipSuccess) e = dalloc(&ctx->d_nrm, 4 * spx, got);
        if (e == hipSuccess) e = dalloc(&ctx->d_wts, spx, got);
        ctx->bytes += got;
        if (e != hipSuccess) return ctx->hip_fail("sgm_set_point_params: allocation", e);
    }
    ctx->geom = *geometry;
    ctx->pparams = *params;
    ctx->have_points = true;
    return SGM_OK;
}

int sgm_compute_frame_raw(sgm_context* ctx, const uint8_t* raw_left, const uint8_t* raw_right, ptrdiff_t raw_stride,
                          int raw_w, int raw_h, const sgm_frame* out, uint8_t* rect_left, uint8_t* rect_right) {
    if (!ctx) {
        g_last_error = "ctx is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = check_frame(ctx); rc != SGM_OK) return rc;
    if (int rc = check_raw(ctx, raw_w, raw_h); rc != SGM_OK) return rc;
    if (!raw_left || !raw_right || raw_stride < raw_w) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "null image or stride < width");
    const int w = ctx->map_w / 2, h = ctx->map_h / 2;
    if (int rc = check_stereo_size(ctx, w, h); rc != SGM_OK) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = ctx->stream;
    if (const hipError_t oe = order_after_last(ctx, s); oe != hipSuccess) return ctx->hip_fail("stream order", oe);
    const size_t npx = (size_t)w * h, rpx = (size_t)raw_w * raw_h;
    uint8_t* rl = ctx->d_img;
    uint8_t* rr = ctx->d_img + npx;
    Stages* st = next_stages(ctx);
    hipError_t e = hipMemcpy2DAsync(ctx->d_raw, raw_w, raw_left, raw_stride, raw_w, raw_h, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpy2DAsync(ctx->d_raw + rpx, raw_w, raw_right, raw_stride, raw_w, raw_h, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return ctx->hip_fail("sgm_compute_frame_raw: upload", e);
    const sgm_frame dev{ctx->d_pts, ctx->d_nrm, ctx->d_wts};
    if ((e = enqueue_frame(ctx, ctx->d_raw, ctx->d_raw + rpx, raw_w, rl, rr, dev, s, st)) != hipSuccess)
        return ctx->hip_fail("sgm_compute_frame_raw: launch", e);
    struct Copy {
        void* host;
        const void* dev;
        size_t bytes;
    };
    const Copy copies[] = {{out ? out->points_xyzw : nullptr, ctx->d_pts, npx * 16},
                           {out ? out->normals_xyzw : nullptr, ctx->d_nrm, npx * 16},
                           {out ? out->weights : nullptr, ctx->d_wts, npx * 4},
                           {rect_left, rl, npx},
                           {rect_right, rr, npx}};
    for (const Copy& c : copies)
        if (c.host && (e = hipMemcpyAsync(c.host, c.dev, c.bytes, hipMemcpyDeviceToHost, s)) != hipSuccess)
            return ctx->hip_fail("sgm_compute_frame_raw: download", e);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return ctx->hip_fail("sgm_compute_frame_raw: execution", e);
    ctx->last_stream = s;
    ctx->pending_wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    read_timings(ctx);
    return SGM_OK;
}

int sgm_compute_frame_raw_device(sgm_context* ctx, const uint8_t* d_raw_left, const uint8_t* d_raw_right,
                                 ptrdiff_t raw_pitch, int raw_w, int raw_h, const sgm_frame* d_out, uint8_t* d_rect_left,
                                 uint8_t* d_rect_right, void* stream) {
    if (!ctx) {
        g_last_error = "ctx is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = check_frame(ctx); rc != SGM_OK) return rc;
    if (int rc = check_raw(ctx, raw_w, raw_h); rc != SGM_OK) return rc;
    if (!d_raw_left || !d_raw_right || raw_pitch < raw_w) return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "null image or pitch < width");
    const int w = ctx->map_w / 2, h = ctx->map_h / 2;
    if (int rc = check_stereo_size(ctx, w, h); rc != SGM_OK) return rc;
    sgm_frame dev{ctx->d_pts, ctx->d_nrm, ctx->d_wts};
    if (d_out) {
        if (d_out->points_xyzw) dev.points_xyzw = d_out->points_xyzw;
        if (d_out->normals_xyzw) dev.normals_xyzw = d_out->normals_xyzw;
        if (d_out->weights) dev.weights = d_out->weights;
    }
    if (!aligned16(dev.points_xyzw) || !aligned16(dev.normals_xyzw))
        return ctx->fail(SGM_ERR_INVALID_ARGUMENT, "point and normal arrays must be 16-byte aligned");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    if (const hipError_t oe = order_after_last(ctx, s); oe != hipSuccess) return ctx->hip_fail("stream order", oe);
    uint8_t* rl = d_rect_left ? d_rect_left : ctx->d_img;
    uint8_t* rr = d_rect_right ? d_rect_right : ctx->d_img + (size_t)w * h;
    const hipError_t e = enqueue_frame(ctx, d_raw_left, d_raw_right, raw_pitch, rl, rr, dev, s, next_stages(ctx));
    if (e != hipSuccess) return ctx->hip_fail("sgm_compute_frame_raw_device: launch", e);
    ctx->last_stream = s;
    ctx->pending_wall_ms = 0;
    return SGM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------ batches
//
// A batch runs as one hipGraph: a census launch over every pair; per lane group (pairs sharing a
// lane layout, optionally split in two by work) a cost launch, a path launch and a row launch, each
// covering all the group's pairs through a flat block -> (pair, block) table; groups run as
// parallel graph branches; one median launch closes the graph. How it is laid out (cost mode,
// scanlines per path block, branches) is the plan; sgm_batch_create times candidate plans on the
// batch itself and keeps the fastest (cached per batch signature for the process).

namespace {

struct BatchPlan {
    int cost_volume = 0;  // 0: recompute costs in the path kernel, 1: cost volume
    int path_lines = 16;  // scanlines per path-kernel block
    int branches = 1;     // 0: one stream, 1: a branch per lane class, 2: two per class
    int path_group = 16;  // lanes per scanline: 16, or 32 (at most 8 lines) for classes that allow it
    int sweep_band = 0;   // > 0: band sweeps of this many rows (sweep_kernels.hip); 0: lines only
    int sweep_tile = 64;  // band sweeps: columns per block
};

// Lanes per scanline of a class under a plan: 32 only where the lane layout allows it.
int class_group(const BatchPlan& pl, int dpl) { return paths_group_ok(dpl, pl.path_group, pl.path_lines) ? pl.path_group : 16; }

// A lane class's sweep block shape under a plan: the plan's tile and band, narrowed (tile first) until
// the kernel takes it with the block's LDS in half a CU's, so that two blocks share a CU.
constexpr size_t kSweepLdsLimit = 80 * 1024;
void sweep_shape(const BatchPlan& pl, int dpl, int& tile, int& band) {
    tile = pl.sweep_tile;
    band = pl.sweep_band;
    while (!sweep_shape_ok(dpl, tile, band, kSweepLdsLimit) && tile > 16) tile -= 16;
    while (!sweep_shape_ok(dpl, tile, band, kSweepLdsLimit) && band > 1) band = (band + 1) / 2;
}

std::mutex g_plan_mu;
std::map<std::string, BatchPlan> g_plan_cache;

}  // namespace

struct sgm_batch {
    sgm_params p{};
    SgmConsts consts{};
    int n = 0;
    int ndirs = 4;
    std::vector<PairDesc> desc;  // host copies; cvol points at each pair's cost volume
    uint8_t* d_scratch = nullptr;
    size_t bytes = 0;
    hipStream_t stream = nullptr;
    std::vector<hipStream_t> branch;  // extra streams for parallel branches
    std::vector<hipEvent_t> events;
    bool sweep_ok = false;            // band sweeps possible: 8 paths, byte path costs, cost volumes
    std::vector<hipStream_t> sweep_streams;  // one per lane class: its horizontal line launch
    std::vector<hipEvent_t> sweep_events;    // two per lane class (fork, join)
    BatchPlan plan;
    std::vector<void*> plan_allocs;  // device tables of the current plan's graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t order_ev = nullptr;  // orders a run on another stream after the previous run's
    sgm_batch_plan info{};
    std::string signature;
    hipStream_t last_stream = nullptr;
    std::string err;
    int fail(int code, std::string m) {
        err = std::move(m);
        return code;
    }
};

namespace {

// Uploads a descriptor array (or inlines a single pair) and its flat-block prefix table.
hipError_t make_table(sgm_batch* b, const std::vector<PairDesc>& ds, const std::vector<int>& blocks, LaunchTable& t) {
    t = LaunchTable{};
    t.n = (int)ds.size();
    std::vector<int32_t> first(ds.size() + 1, 0);
    for (size_t i = 0; i < ds.size(); ++i) first[i + 1] = first[i] + blocks[i];
    t.total_blocks = first.back();
    if (ds.size() == 1) {
        t.one = ds[0];
        return hipSuccess;
    }
    PairDesc* dd = nullptr;
    int32_t* df = nullptr;
    size_t unused = 0;
    hipError_t e = dalloc(&dd, ds.size(), unused);
    if (e == hipSuccess) b->plan_allocs.push_back(dd);
    if (e == hipSuccess) e = dalloc(&df, first.size(), unused);
    if (e == hipSuccess) b->plan_allocs.push_back(df);
    if (e == hipSuccess) e = hipMemcpy(dd, ds.data(), ds.size() * sizeof(PairDesc), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(df, first.data(), first.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    t.many = dd;
    t.first_block = df;
    return e;
}

struct Group {
    int dpl = 0;
    bool padded = false;
    int max_w = 1;
    std::vector<PairDesc> ds;
    LaunchTable cost, paths, rows;
    // Band sweeps: the tables of both sweeps, the block shape and the number of band launches.
    LaunchTable sweep;
    int tile = 0, band = 0, nbands = 0;
};

// Band-sweep plans: a pair's path-cost buffer holds one record per pixel, [S_down[dp] u16,
// S_up[dp] u16, L0[dp] u8, L1[dp] u8] (6 dp bytes, within the 8 dp of the line schedule's layout); the
// line kernel writes only directions 0 and 1, into the record's last 2 dp bytes.
PairDesc sweep_desc(PairDesc d) {
    d.lvol_px_stride = (int64_t)6 * d.dp;
    return d;
}
PairDesc horizontal_desc(PairDesc d) {
    d.lvol = d.lvol + (size_t)4 * d.dp;
    d.lvol_px_stride = (int64_t)6 * d.dp;
    d.ndirs = 2;
    fill_lines(d);
    return d;
}

void drop_plan(sgm_batch* b) {
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    for (void* q : b->plan_allocs) dfree(q);
    b->plan_allocs.clear();
}

// Builds and instantiates the graph of `plan`.
hipError_t build_plan(sgm_batch* b, const BatchPlan& plan) {
    drop_plan(b);
    b->plan = plan;
    std::vector<PairDesc> all = b->desc;
    for (PairDesc& d : all)
        if (!plan.cost_volume) d.cvol = nullptr;
    // Lane classes, each split in `parts` chunks of about equal work.
    std::map<std::pair<int, bool>, std::vector<int>> cls;
    for (int i = 0; i < b->n; ++i) cls[{all[(size_t)i].dpl, all[(size_t)i].dp != all[(size_t)i].nd}].push_back(i);
    std::vector<Group> groups;
    const int parts = plan.branches == 2 ? 2 : 1;
    for (auto& [key, idx] : cls) {
        // Longest work first inside a launch: its scanlines start earliest.
        std::sort(idx.begin(), idx.end(), [&](int x, int y) {
            return (int64_t)all[(size_t)x].w * all[(size_t)x].h * all[(size_t)x].nd > (int64_t)all[(size_t)y].w * all[(size_t)y].h * all[(size_t)y].nd;
        });
        const int k = std::min<int>(parts, (int)idx.size());
        std::vector<Group> gs((size_t)k);
        std::vector<int64_t> work((size_t)k, 0);
        for (int i : idx) {  // greedy: next pair to the lightest chunk
            const size_t g = (size_t)(std::min_element(work.begin(), work.end()) - work.begin());
            work[g] += (int64_t)all[(size_t)i].w * all[(size_t)i].h * all[(size_t)i].nd;
            gs[g].ds.push_back(all[(size_t)i]);
        }
        for (Group& g : gs) {
            g.dpl = key.first;
            g.padded = key.second;
            for (const PairDesc& d : g.ds) g.max_w = std::max(g.max_w, d.w);
            groups.push_back(std::move(g));
        }
    }
    hipError_t e = hipSuccess;
    const bool sweeps = plan.sweep_band > 0;
    if (sweeps && (!b->sweep_ok || !plan.cost_volume)) return hipErrorInvalidValue;
    for (Group& g : groups) {
        std::vector<int> cb, pb, rb, sb;
        std::vector<PairDesc> hds, sds;
        if (sweeps) sweep_shape(plan, g.dpl, g.tile, g.band);
        for (const PairDesc& d : g.ds) {
            cb.push_back(cost8_blocks(d.w, d.h));
            rb.push_back(row_blocks(d));
            if (sweeps) {
                hds.push_back(horizontal_desc(d));
                sds.push_back(sweep_desc(d));
                pb.push_back(path_blocks(hds.back(), plan.path_lines));
                sb.push_back(sweep_blocks(d.w, g.tile));
                g.nbands = std::max(g.nbands, (d.h + g.band - 1) / g.band);
            } else {
                pb.push_back(path_blocks(d, plan.path_lines));
            }
        }
        if (e == hipSuccess && plan.cost_volume) e = make_table(b, g.ds, cb, g.cost);
        if (e == hipSuccess) e = make_table(b, sweeps ? hds : g.ds, pb, g.paths);
        g.paths.cost_volume = plan.cost_volume ? 1 : 0;  // the path kernel reads the costs the cost launch wrote
        if (e == hipSuccess) e = make_table(b, sweeps ? sds : g.ds, rb, g.rows);
        if (e == hipSuccess && sweeps) e = make_table(b, sds, sb, g.sweep);
    }
    LaunchTable census{}, median{};
    std::vector<int> blocks;
    for (const PairDesc& d : all) blocks.push_back(census_blocks(d.w, d.h));
    if (e == hipSuccess) e = make_table(b, all, blocks, census);
    std::vector<PairDesc> med;
    blocks.clear();
    for (const PairDesc& d : all)
        if (d.median) {
            med.push_back(d);
            blocks.push_back(median_blocks(d.w, d.h));
        }
    if (e == hipSuccess && !med.empty()) e = make_table(b, med, blocks, median);
    // Streams and events for the branches.
    const size_t nbr = plan.branches == 0 ? 1 : groups.size();
    while (e == hipSuccess && b->branch.size() + 1 < nbr) {
        hipStream_t st = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) b->branch.push_back(st);
    }
    while (e == hipSuccess && b->events.size() < nbr + 1) {
        hipEvent_t ev = nullptr;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) b->events.push_back(ev);
    }
    // Band sweeps: the down and up sweeps of each class run beside its horizontal line launch.
    while (e == hipSuccess && sweeps && b->sweep_streams.size() < groups.size()) {
        hipStream_t st = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) b->sweep_streams.push_back(st);
    }
    while (e == hipSuccess && sweeps && b->sweep_events.size() < 2 * groups.size()) {
        hipEvent_t ev = nullptr;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) b->sweep_events.push_back(ev);
    }
    if (e != hipSuccess) return e;
    hipStream_t s = b->stream;
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t le = launch_census(census, b->consts, s);
    if (le == hipSuccess && nbr > 1) {
        le = hipEventRecord(b->events[0], s);
        for (size_t i = 1; le == hipSuccess && i < nbr; ++i) le = hipStreamWaitEvent(b->branch[i - 1], b->events[0], 0);
    }
    for (size_t gi = 0; le == hipSuccess && gi < groups.size(); ++gi) {
        const Group& g = groups[gi];
        hipStream_t gs = (nbr == 1 || gi == 0) ? s : b->branch[gi - 1];
        if (plan.cost_volume) le = launch_cost8(g.dpl, g.cost, b->consts, gs);
        if (sweeps) {
            // The class's band sweeps (down and up in each launch) stay on the class's stream; its
            // horizontal line launch runs beside them on a stream of its own.
            hipStream_t hs = b->sweep_streams[gi];
            hipEvent_t* ev = &b->sweep_events[2 * gi];
            if (le == hipSuccess) le = hipEventRecord(ev[0], gs);
            if (le == hipSuccess) le = hipStreamWaitEvent(hs, ev[0], 0);
            if (le == hipSuccess)
                le = launch_paths(g.dpl, g.padded, g.paths, b->consts, plan.path_lines, class_group(plan, g.dpl), hs);
            for (int band = 0; le == hipSuccess && band < g.nbands; ++band)
                le = launch_sweep(g.dpl, g.padded, g.sweep, b->consts, g.tile, g.band, band, gs);
            if (le == hipSuccess) le = hipEventRecord(ev[1], hs);
            if (le == hipSuccess) le = hipStreamWaitEvent(gs, ev[1], 0);
            if (le == hipSuccess) le = launch_rows_sweep(g.dpl, g.padded, g.rows, b->consts, g.max_w, gs);
        } else {
            if (le == hipSuccess) le = launch_paths(g.dpl, g.padded, g.paths, b->consts, plan.path_lines, class_group(plan, g.dpl), gs);
            if (le == hipSuccess) le = launch_rows(g.dpl, g.padded, g.rows, b->consts, g.max_w, gs);
        }
    }
    for (size_t i = 1; le == hipSuccess && nbr > 1 && i < nbr; ++i) {
        le = hipEventRecord(b->events[i], b->branch[i - 1]);
        if (le == hipSuccess) le = hipStreamWaitEvent(s, b->events[i], 0);
    }
    if (le == hipSuccess && !med.empty()) le = launch_median(median, s);
    e = hipStreamEndCapture(s, &b->graph);
    if (le != hipSuccess) return le;
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    return e;
}

// Mean time of one replay of the current plan's graph (after one warm-up replay).
hipError_t time_plan(sgm_batch* b, int reps, float* ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipGraphLaunch(b->exec, b->stream);
    if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
    for (int i = 0; e == hipSuccess && i < reps; ++i) e = hipGraphLaunch(b->exec, b->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    if (e == hipSuccess) *ms /= (float)reps;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return e;
}

// The plan cache's key: the parameters that select kernel variants (path count, census window,
// slant steps) and the pairs' shapes bucketed -- width and height rounded up to 64, the lane layout
// (disparities per lane and whether they are padded), the median flag -- in sorted order. Batches of
// nearly the same shapes share a plan, which is safe because a plan never changes a result.
std::string batch_signature(const sgm_params& p, int n, const sgm_pair* pairs) {
    std::vector<std::string> items;
    for (int i = 0; i < n; ++i) {
        const int dpl = dpl_for(pairs[i].num_disparities);
        items.push_back(std::to_string((pairs[i].width + 63) / 64) + "x" + std::to_string((pairs[i].height + 63) / 64) + "d" +
                        std::to_string(dpl) + (dpl * kGroup != pairs[i].num_disparities ? "p" : "") +
                        (pairs[i].out.disparity_median ? "m" : ""));
    }
    std::sort(items.begin(), items.end());
    std::string k = "v3:" + std::to_string(p.eight_paths ? 8 : 4) + ":" + std::to_string(p.census_radius_x) + "x" +
                    std::to_string(p.census_radius_y) + ":s" + std::to_string(p.slant_steps ? 1 : 0);
    for (const std::string& it : items) k += ";" + it;
    return k;
}

}  // namespace

extern "C" {

void sgm_batch_destroy(sgm_batch* b) {
    if (!b) return;
    // A run on a caller's stream may still use the graph, its tables and the scratch.
    if (b->last_stream && b->last_stream != b->stream) (void)hipStreamSynchronize(b->last_stream);
    drop_plan(b);
    dfree(b->d_scratch);
    if (b->order_ev) (void)hipEventDestroy(b->order_ev);
    for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
    for (hipStream_t st : b->branch) (void)hipStreamDestroy(st);
    for (hipEvent_t e : b->sweep_events) (void)hipEventDestroy(e);
    for (hipStream_t st : b->sweep_streams) (void)hipStreamDestroy(st);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* sgm_batch_last_error(const sgm_batch* b) { return b ? b->err.c_str() : g_last_error.c_str(); }
size_t sgm_batch_device_bytes(const sgm_batch* b) { return b ? b->bytes : 0; }

int sgm_batch_get_plan(const sgm_batch* b, sgm_batch_plan* out) {
    if (!b || !out) return SGM_ERR_INVALID_ARGUMENT;
    *out = b->info;
    return SGM_OK;
}

int sgm_batch_create_tuned(const sgm_params* params, int n, const sgm_pair* pairs, const sgm_batch_tuning* tuning,
                           sgm_batch** out) {
    if (!out) {
        g_last_error = "out is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    std::string err;
    if (int rc = validate_params(params, err); rc != SGM_OK) {
        g_last_error = err;
        return rc;
    }
    if (n < 1 || n > kMaxBatch || !pairs) {
        g_last_error = "batch size must be in [1, " + std::to_string(kMaxBatch) + "]";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    sgm_batch_tuning tn{-1, -1, -1, -1, -1, -1};
    if (tuning) tn = *tuning;
    if ((tn.cost_volume < -1 || tn.cost_volume > 1) || (tn.path_lines != -1 && tn.path_lines != 4 && tn.path_lines != 8 && tn.path_lines != 16) ||
        (tn.branches < -1 || tn.branches > 2) || (tn.path_group != -1 && tn.path_group != 16 && tn.path_group != 32) ||
        (tn.path_group == 32 && tn.path_lines == 16) || tn.sweep_band < -1 || tn.sweep_band > 1024 ||
        (tn.sweep_tile != -1 && (tn.sweep_tile < 16 || tn.sweep_tile > 64 || tn.sweep_tile % 16 != 0))) {
        g_last_error = "tuning: cost_volume in {-1,0,1}, path_lines in {-1,4,8,16}, branches in {-1,0,1,2}, "
                       "path_group in {-1,16,32} (32 with at most 8 lines), sweep_band in [-1, 1024], "
                       "sweep_tile in {-1,16,32,48,64}";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    for (int i = 0; i < n; ++i) {
        const sgm_pair& q = pairs[i];
        sgm_params pp = *params;
        pp.min_disparity = q.min_disparity;
        pp.num_disparities = q.num_disparities;
        if (validate_params(&pp, err) != SGM_OK || q.width < 1 || q.height < 1 || q.width > kMaxWidth ||
            q.pitch < q.width || !q.left || !q.right || !q.out.disparity) {
            g_last_error = "pair " + std::to_string(i) + ": " + (err.empty() ? "bad size, pointer or pitch" : err);
            return SGM_ERR_INVALID_ARGUMENT;
        }
        if (q.out.census_left || q.out.census_right || q.out.cost_volume) {
            g_last_error = "pair " + std::to_string(i) + ": census and cost-volume outputs are only available from sgm_compute";
            return SGM_ERR_UNSUPPORTED;
        }
        if (census_words(q.width, q.height, q.min_disparity, q.num_disparities) * 8 > kMaxViewBytes) {
            g_last_error = "pair " + std::to_string(i) + ": the census words of one image exceed 4 GiB - 1";
            return SGM_ERR_INVALID_ARGUMENT;
        }
        if (int rc = check_row_lds(q.width, err); rc != SGM_OK) {
            g_last_error = "pair " + std::to_string(i) + ": " + err;
            return rc;
        }
    }
    // The cost-volume mode needs every pair's volume within one buffer view.
    bool cv_fits = true;
    for (int i = 0; i < n; ++i)
        cv_fits = cv_fits && (size_t)pairs[i].width * pairs[i].height * dpl_for(pairs[i].num_disparities) * kGroup <= kMaxViewBytes;
    if (tn.cost_volume == 1 && !cv_fits) {
        g_last_error = "tuning: a pair's cost volume exceeds 4 GiB - 1";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    const SgmConsts consts = make_consts(*params);
    const bool sweep_ok = params->eight_paths && !consts.wide_costs && cv_fits;
    if (tn.sweep_band > 0 && (!sweep_ok || tn.cost_volume == 0)) {
        g_last_error = "tuning: band sweeps need 8 paths, byte path costs and the cost volume";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    auto* b = new sgm_batch;
    b->p = *params;
    b->consts = consts;
    b->n = n;
    b->ndirs = params->eight_paths ? 8 : 4;
    b->sweep_ok = sweep_ok;
    // Scratch per pair: census words (2 images, padded rows), u8 cost volume, u8 path costs, and for
    // band sweeps the seed rows (two sweeps x two parities x three directions of W x dp bytes).
    std::vector<size_t> off_census(n), off_cvol(n), off_lvol(n), off_seed(n);
    size_t total = 0;
    for (int i = 0; i < n; ++i) {
        const size_t npx = (size_t)pairs[i].width * pairs[i].height;
        const size_t dp = (size_t)dpl_for(pairs[i].num_disparities) * kGroup;
        off_census[i] = total;
        total += (2 * census_words(pairs[i].width, pairs[i].height, pairs[i].min_disparity, pairs[i].num_disparities) * 8 +
                  npx * edge_bytes_per_px(b->consts.wide_costs != 0) + 255) & ~(size_t)255;  // census words, then penalties
        off_cvol[i] = total;
        if (cv_fits) total += (npx * dp + 255) & ~(size_t)255;
        off_lvol[i] = total;
        total += ((size_t)b->ndirs * npx * dp * (b->consts.wide_costs ? 2 : 1) + 255) & ~(size_t)255;
        off_seed[i] = total;
        if (sweep_ok) total += ((size_t)12 * pairs[i].width * dp + 255) & ~(size_t)255;
    }
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = dalloc(&b->d_scratch, total, b->bytes);
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_batch_create: allocation", e);
        sgm_batch_destroy(b);
        return SGM_ERR_IO;
    }
    for (int i = 0; i < n; ++i) {
        const sgm_pair& q = pairs[i];
        PairDesc d{};
        d.left = q.left;
        d.right = q.right;
        d.pitch = q.pitch;
        d.census_l = reinterpret_cast<uint64_t*>(b->d_scratch + off_census[i]);
        d.census_r = d.census_l + census_words(q.width, q.height, q.min_disparity, q.num_disparities);
        d.edge = reinterpret_cast<uint32_t*>(d.census_r + census_words(q.width, q.height, q.min_disparity, q.num_disparities));
        d.cvol = cv_fits ? b->d_scratch + off_cvol[i] : nullptr;
        d.lvol = b->d_scratch + off_lvol[i];
        d.seed = sweep_ok ? b->d_scratch + off_seed[i] : nullptr;
        set_geometry(d, q.width, q.height, q.min_disparity, q.num_disparities, b->ndirs, b->consts.wide_costs ? 2 : 1);
        d.disp = q.out.disparity;
        d.median = q.out.disparity_median;
        d.conf = q.out.confidence;
        d.dint = q.out.disparity_int;
        d.status = q.out.status;
        d.rdisp = q.out.right_disparity;
        d.best = q.out.best_cost;
        d.second = q.out.second_cost;
        d.sum_volume = q.out.sum_volume;
        d.cost_volume = nullptr;
        b->desc.push_back(d);
    }
    // Candidate plans: every combination of the knobs left open.
    std::vector<BatchPlan> cands;
    for (int cv : {0, 1})
        for (int pl : {4, 8, 16})
            for (int br : {0, 1, 2})
                for (int pg : {16, 32}) {
                    if ((tn.cost_volume >= 0 && cv != tn.cost_volume) || (tn.path_lines > 0 && pl != tn.path_lines) ||
                        (tn.branches >= 0 && br != tn.branches) || (tn.path_group > 0 && pg != tn.path_group))
                        continue;
                    if (tn.branches < 0 && br == 0) continue;  // one stream only when asked for
                    if (pg == 32 && pl == 16) continue;           // 32-lane lines: at most 8 per block
                    if (cv == 1 && !cv_fits) continue;
                    if (tn.sweep_band <= 0) cands.push_back({cv, pl, br, pg, 0, 64});
                    // Band sweeps: a few band heights and tiles, beside the horizontal line launch of
                    // each knob setting that the caller fixed (8 lines of 16 lanes otherwise).
                    if (tn.sweep_band == 0 || !sweep_ok || cv != 1) continue;
                    if ((tn.path_lines < 0 && pl != 8) || (tn.path_group < 0 && pg != 16) || (tn.branches < 0 && br != 1))
                        continue;
                    const std::vector<int> bands = tn.sweep_band > 0 ? std::vector<int>{tn.sweep_band} : std::vector<int>{16, 32, 64};
                    const std::vector<int> tiles = tn.sweep_tile > 0 ? std::vector<int>{tn.sweep_tile} : std::vector<int>{32, 64};
                    for (int band : bands)
                        for (int tile : tiles) cands.push_back({cv, pl, br, pg, band, tile});
                }
    if (cands.empty()) {
        g_last_error = "tuning: no plan satisfies the fixed knobs";
        sgm_batch_destroy(b);
        return SGM_ERR_INVALID_ARGUMENT;
    }
    b->signature = batch_signature(*params, n, pairs);
    BatchPlan best = cands.front();
    float best_ms = 0, worst_ms = 0;
    bool cached = false;
    // Only a search over every knob is cached, and a cached plan is taken only where it honours the
    // knobs this call fixed.
    const bool all_open = tn.cost_volume < 0 && tn.path_lines < 0 && tn.branches < 0 && tn.path_group < 0 &&
                          tn.sweep_band < 0 && tn.sweep_tile < 0;
    if (cands.size() > 1) {
        std::lock_guard lk(g_plan_mu);
        auto it = g_plan_cache.find(b->signature);
        if (it != g_plan_cache.end() && std::any_of(cands.begin(), cands.end(), [&](const BatchPlan& c) {
                return c.cost_volume == it->second.cost_volume && c.path_lines == it->second.path_lines &&
                       c.branches == it->second.branches && c.path_group == it->second.path_group &&
                       c.sweep_band == it->second.sweep_band && (c.sweep_band == 0 || c.sweep_tile == it->second.sweep_tile);
            })) {
            best = it->second;
            cached = true;
        }
    }
    int timed = 0;
    if (!cached && cands.size() > 1) {
        for (const BatchPlan& pl : cands) {
            float ms = 0;
            e = build_plan(b, pl);
            if (e == hipSuccess) e = time_plan(b, 2, &ms);
            if (e != hipSuccess) break;
            ++timed;
            if (timed == 1 || ms < best_ms) {
                best_ms = ms;
                best = pl;
            }
            worst_ms = std::max(worst_ms, ms);
        }
        if (e == hipSuccess && all_open) {
            std::lock_guard lk(g_plan_mu);
            g_plan_cache[b->signature] = best;
        }
    }
    if (e == hipSuccess) e = build_plan(b, best);
    if (e != hipSuccess) {
        g_last_error = hip_msg("sgm_batch_create: plan/graph", e);
        sgm_batch_destroy(b);
        return SGM_ERR_IO;
    }
    b->info.cost_volume = best.cost_volume;
    b->info.path_lines = best.path_lines;
    b->info.path_group = best.path_group;
    b->info.branches = best.branches;
    b->info.sweep_band = best.sweep_band;
    b->info.sweep_tile = best.sweep_tile;
    b->info.candidates_timed = timed;
    b->info.best_ms = best_ms;
    b->info.worst_ms = worst_ms;
    b->info.from_cache = cached ? 1 : 0;
    *out = b;
    return SGM_OK;
}

int sgm_tune_cache_load(const char* path, int* entries) {
    if (!path) {
        g_last_error = "path is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    FILE* f = std::fopen(path, "r");
    if (!f) {
        g_last_error = std::string("sgm_tune_cache_load: cannot open ") + path;
        return SGM_ERR_NOT_FOUND;
    }
    // One entry a line: key, a tab, then cost_volume path_lines branches path_group sweep_band sweep_tile.
    std::map<std::string, BatchPlan> got;
    char line[8192];
    int rc = SGM_OK;
    while (std::fgets(line, sizeof line, f)) {
        std::string ln(line);
        while (!ln.empty() && (ln.back() == '\n' || ln.back() == '\r')) ln.pop_back();
        if (ln.empty()) continue;
        const size_t tab = ln.find('\t');
        BatchPlan pl;
        if (tab == std::string::npos || tab == 0 ||
            std::sscanf(ln.c_str() + tab + 1, "%d %d %d %d %d %d", &pl.cost_volume, &pl.path_lines, &pl.branches, &pl.path_group,
                        &pl.sweep_band, &pl.sweep_tile) != 6 ||
            (pl.cost_volume != 0 && pl.cost_volume != 1) || (pl.path_lines != 4 && pl.path_lines != 8 && pl.path_lines != 16) ||
            pl.branches < 0 || pl.branches > 2 || (pl.path_group != 16 && pl.path_group != 32) || pl.sweep_band < 0 ||
            pl.sweep_band > 1024 || (pl.sweep_band > 0 && pl.cost_volume != 1) || pl.sweep_tile < 16 || pl.sweep_tile > 64 ||
            pl.sweep_tile % 16 != 0) {
            g_last_error = "sgm_tune_cache_load: malformed line: " + ln.substr(0, 200);
            rc = SGM_ERR_PROTOCOL;
            break;
        }
        got[ln.substr(0, tab)] = pl;
    }
    std::fclose(f);
    if (rc != SGM_OK) return rc;
    std::lock_guard lk(g_plan_mu);
    for (auto& [k, v] : got) g_plan_cache[k] = v;
    if (entries) *entries = (int)got.size();
    return SGM_OK;
}

int sgm_tune_cache_save(const char* path, int* entries) {
    if (!path) {
        g_last_error = "path is NULL";
        return SGM_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard lk(g_plan_mu);
    FILE* f = std::fopen(path, "w");
    if (!f) {
        g_last_error = std::string("sgm_tune_cache_save: cannot open ") + path;
        return SGM_ERR_IO;
    }
    for (const auto& [k, v] : g_plan_cache)
        std::fprintf(f, "%s\t%d %d %d %d %d %d\n", k.c_str(), v.cost_volume, v.path_lines, v.branches, v.path_group,
                     v.sweep_band, v.sweep_tile);
    const bool ok = std::fclose(f) == 0;
    if (!ok) {
        g_last_error = std::string("sgm_tune_cache_save: write failed: ") + path;
        return SGM_ERR_IO;
    }
    if (entries) *entries = (int)g_plan_cache.size();
    return SGM_OK;
}

int sgm_batch_create(const sgm_params* params, int n, const sgm_pair* pairs, sgm_batch** out) {
    return sgm_batch_create_tuned(params, n, pairs, nullptr, out);
}

// Runs of one batch share its scratch, so a run on another stream than the previous run's waits for
// that stream's work first.
static hipError_t batch_order(sgm_batch* b, hipStream_t s) {
    if (!b->last_stream || b->last_stream == s) return hipSuccess;
    hipError_t e = hipSuccess;
    if (!b->order_ev) e = hipEventCreateWithFlags(&b->order_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(b->order_ev, b->last_stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, b->order_ev, 0);
    return e;
}

int sgm_batch_run(sgm_batch* b, void* stream) {
    if (!b) return SGM_ERR_INVALID_ARGUMENT;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : b->stream;
    if (const hipError_t oe = batch_order(b, s); oe != hipSuccess) return b->fail(SGM_ERR_IO, hip_msg("stream order", oe));
    const hipError_t e = hipGraphLaunch(b->exec, s);
    if (e != hipSuccess) return b->fail(SGM_ERR_IO, hip_msg("sgm_batch_run", e));
    b->last_stream = s;
    return SGM_OK;
}

int sgm_batch_run_plain(sgm_batch* b, void* stream) {
    if (!b) return SGM_ERR_INVALID_ARGUMENT;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : b->stream;
    if (const hipError_t oe = batch_order(b, s); oe != hipSuccess) return b->fail(SGM_ERR_IO, hip_msg("stream order", oe));
    for (PairDesc d : b->desc) {
        // The same kernels and knobs as the graph, one pair after another: only batching differs.
        if (!b->plan.cost_volume) d.cvol = nullptr;
        const hipError_t e = enqueue_pair(d, b->consts, d.dp != d.nd, s, nullptr, b->plan.path_lines,
                                          class_group(b->plan, d.dpl));
        if (e != hipSuccess) return b->fail(SGM_ERR_IO, hip_msg("sgm_batch_run_plain", e));
    }
    b->last_stream = s;
    return SGM_OK;
}

int sgm_batch_synchronize(sgm_batch* b) {
    if (!b) return SGM_ERR_INVALID_ARGUMENT;
    const hipError_t e = hipStreamSynchronize(b->last_stream ? b->last_stream : b->stream);
    if (e != hipSuccess) return b->fail(SGM_ERR_IO, hip_msg("sgm_batch_synchronize", e));
    return SGM_OK;
}

}  // extern "C"
