/*
 * sgm_hip.h -- C ABI of the MI355X (gfx950) SGM stereo matcher.
 *
 * Drop-in replacement for the GPU stereo backend of the Einstar scanner software. The reference
 * binds its GPU matcher as the C++ class depth_metal::MetalStereo
 * (libs/depth_metal/include/einstar/depth_metal/metal_stereo.hpp:67-107), whose SGM core runs the
 * algorithm of depth::sgm_disparity (libs/depth/include/einstar/depth/stereo.hpp:113-115,
 * libs/depth/src/stereo.cpp:361-463). This header is the C boundary that class's replacement
 * (include/einstar_hip/hip_stereo.hpp) and foreign callers (ctypes, cgo, JNI) bind: plain
 * pointers, sizes and status codes; no C++ or HIP types cross it (streams travel as void*).
 *
 * Threading: a context is not re-entrant, like MetalStereo (the reference serialises it with a
 * mutex, libs/pipeline/src/stereo_frontend.cpp:307). Separate contexts are independent. The
 * server functions are safe to call from any thread.
 *
 * Streams: every call on one context (or one batch) uses that context's scratch buffers. A call
 * enqueued on another stream than the previous call's is made to wait, on the device, for the
 * previous call's stream, so calls on one context never overlap whatever streams they name.
 * Outputs written by a call on a caller's stream are ready once that stream has reached them.
 */
#ifndef SGM_HIP_H
#define SGM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGM_HIP_ABI_VERSION 2

/* Status codes: the values of einstar::Errc (libs/core/include/einstar/core/error.hpp:8-19). */
enum {
    SGM_OK = 0,
    SGM_ERR_NOT_FOUND = 1,
    SGM_ERR_IO = 2, /* GPU runtime failure (MetalStereo returns Errc::io, metal_stereo.cpp:716-718) */
    SGM_ERR_TIMEOUT = 3,
    SGM_ERR_PROTOCOL = 4,
    SGM_ERR_INVALID_ARGUMENT = 5, /* size or parameter out of range (metal_stereo.cpp:115-120, 235-236) */
    SGM_ERR_BLOCKED = 6,
    SGM_ERR_UNSUPPORTED = 7,
    SGM_ERR_BUSY = 8, /* serving: not admitted (no free slot or HBM) */
    SGM_ERR_DISCONNECTED = 9
};

/* kInvalidDisparity (stereo.hpp:23) and the right-WTA "no match" (stereo.cpp:413). */
#define SGM_INVALID_DISPARITY (-1.0e9f)
#define SGM_NO_MATCH INT32_MIN

/* Per-pixel status bits of the left WTA chain (stereo.cpp:419-461); valid == all three. */
#define SGM_STATUS_WINDOW 1u /* the candidate window [min_k, max_k) is not empty */
#define SGM_STATUS_UNIQUE 2u /* passed the uniqueness test */
#define SGM_STATUS_LR 4u     /* passed the left/right check */
#define SGM_STATUS_VALID 7u

/* depth::SgmParams (stereo.hpp:26-40) plus StereoParams::subpixel (stereo.hpp:87). Booleans are
 * 0/1. Defaults: sgm_params_default(). Limits (those of MetalStereo::create,
 * metal_stereo.cpp:115-120): 1 <= num_disparities <= 512, census window <= 9x7,
 * window cells + p2 <= 255; and 0 <= p1 <= 2^30, p2 >= 0. Path costs are bytes where the largest step
 * penalty allows (cells + max(p2, p1 + 1 if adaptive) <= 255 and p1 <= 32767), u16 otherwise; results
 * are the same either way. Image widths up to 32768 columns. */
typedef struct {
    int32_t min_disparity;
    int32_t num_disparities;
    int32_t census_radius_x;
    int32_t census_radius_y;
    int32_t p1;
    int32_t p2;
    int32_t adaptive_p2;
    int32_t eight_paths;
    int32_t slant_steps;
    float uniqueness;
    int32_t lr_max_diff;
    int32_t subpixel;
} sgm_params;

/* Outputs, each W*H (row pitch = width) unless noted; any may be NULL except where noted.
 * Host pointers for sgm_compute, device pointers for sgm_compute_device and batches. */
typedef struct {
    float* disparity;          /* sgm_disparity() result (stereo.hpp:114); invalid = SGM_INVALID_DISPARITY */
    float* disparity_median;   /* median3(disparity) (stereo.cpp:129-147) */
    float* confidence;         /* 1 where disparity is valid, else 0 (stereo.cpp:515-519) */
    int32_t* disparity_int;    /* min_disparity + best_k, SGM_NO_MATCH for an empty window */
    uint8_t* status;           /* SGM_STATUS_* bits */
    int32_t* right_disparity;  /* right WTA (stereo.cpp:434-447), SGM_NO_MATCH where none */
    uint16_t* best_cost;       /* aggregated cost at best_k, 0xFFFF for an empty window */
    uint16_t* second_cost;     /* best non-adjacent aggregated cost, 0xFFFF if none */
    uint16_t* sum_volume;      /* debug: aggregated costs S, W*H*D, [y][x][d] */
    uint16_t* cost_volume;     /* debug: matching costs C, W*H*D, [y][x][d] */
    uint64_t* census_left;     /* debug: census words */
    uint64_t* census_right;
} sgm_outputs;

/* Per-stage GPU times of the last completed compute (hipEvents on the context's stream), as
 * MetalStereo::last_timings (metal_stereo.hpp:61-65, 97). */
typedef struct {
    double census_ms;
    double paths_ms;
    double rows_ms;   /* sum of paths, WTA, left/right check */
    double median_ms;
    double gpu_ms;    /* first to last kernel */
    double total_ms;  /* host wall time of the call, including copies */
    double rectify_ms; /* rectification (raw calls) and pyramid (compute_disparity), else 0 */
    double refine_ms;  /* compute_disparity: ZNCC refinement of every level and the detail refinement */
    double filter_ms;  /* compute_disparity: occlusion and spike filters, speckle removal, confidence */
    double points_ms;  /* compute_frame: points, normals and weights */
} sgm_timings;

/* Per-stage GPU times summed over every call completed since the last reset (count = calls). */
typedef struct {
    double census_ms;
    double paths_ms;
    double rows_ms;
    double median_ms;
    double gpu_ms;
    int64_t count;
    double rectify_ms;
    double refine_ms;
    double filter_ms;
    double points_ms;
} sgm_stage_totals;

typedef struct sgm_context sgm_context;

/* Fills the reference defaults (stereo.hpp:26-40, subpixel = true). */
void sgm_params_default(sgm_params* p);

/* Replaces MetalStereo::create (metal_stereo.hpp:70-71): validates the parameters and allocates
 * device buffers for pairs up to max_width x max_height on the current HIP device. */
int sgm_create(const sgm_params* params, int max_width, int max_height, sgm_context** out);

/* Replaces ~MetalStereo (metal_stereo.hpp:72): frees every buffer. NULL is ignored. */
void sgm_destroy(sgm_context* ctx);

/* Error text of the last failing call on ctx, or (ctx == NULL) of the last failing sgm_create
 * or server call on this thread. Never NULL. */
const char* sgm_last_error(const sgm_context* ctx);

/* Replaces MetalStereo::compute (metal_stereo.hpp:78) for its SGM stage, i.e. depth::sgm_disparity
 * on rectified images (stereo.hpp:114-115) plus median3: host images with a row stride in bytes,
 * host outputs; blocks until they are written. */
int sgm_compute(sgm_context* ctx, const uint8_t* left, const uint8_t* right, ptrdiff_t stride, int width, int height,
                const sgm_outputs* out);

/* The same on device-resident images and outputs, enqueued on `stream` (a hipStream_t; NULL =
 * the context's own stream) and returning without waiting. out->disparity may be NULL (the
 * context's buffer is used). Timings become available after sgm_synchronize. */
int sgm_compute_device(sgm_context* ctx, const uint8_t* d_left, const uint8_t* d_right, ptrdiff_t pitch, int width,
                       int height, const sgm_outputs* d_out, void* stream);

/* Bytes of device scratch that sgm_compute_device_noalloc has on each GPU (128 MiB). */
#define SGM_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)

/* sgm_compute_device without a context and without allocating: no counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * hosts that forbid hipMalloc). Same inputs, outputs, results and stream semantics as
 * sgm_compute_device (params as sgm_create takes them; d_out may be NULL or leave disparity NULL);
 * census outputs are refused. The call allocates nothing: its scratch (census words, step
 * penalties, path costs, the cost volume where it fits, the disparity map when d_out has none) lives
 * in one static device array of SGM_NOALLOC_SCRATCH_BYTES per GPU, held from when the library's
 * device code is loaded on that GPU (at the latest, the first call there) to the end of the process. Pairs whose scratch exceeds it are refused with
 * SGM_ERR_INVALID_ARGUMENT (e.g. 640x512 with 64 disparities and 4 paths fits, 1280x1024x128 with 8
 * paths does not). Every call on one GPU shares that array: calls whose work overlaps on one GPU
 * (different streams, or different host threads, without ordering between them) race and give wrong
 * results; calls on one stream are ordered. Errors: sgm_last_error(NULL) on the calling thread. */
int sgm_compute_device_noalloc(const sgm_params* params, const uint8_t* d_left, const uint8_t* d_right, ptrdiff_t pitch,
                               int width, int height, const sgm_outputs* d_out, void* stream);
/* SGM_NOALLOC_SCRATCH_BYTES, as built into the library. */
size_t sgm_noalloc_scratch_bytes(void);

/* Replaces MetalStereo::set_rectification (metal_stereo.hpp:74-75, metal_stereo.cpp:203-231): uploads
 * the left and right calib::RemapTable maps (libs/calib/include/einstar/calib/rectify.hpp:40-43;
 * map_x and map_y as row-major map_w x map_h host arrays of source pixel coordinates) once; they stay
 * resident on the device until the next call or sgm_destroy. The stereo size becomes
 * (map_w / 2) x (map_h / 2), which must fit the context ("remap resolution must be twice the stereo
 * resolution"); raw frames are raw_w x raw_h. */
int sgm_set_rectification(sgm_context* ctx, const float* left_map_x, const float* left_map_y, const float* right_map_x,
                          const float* right_map_y, int map_w, int map_h, int raw_w, int raw_h);

/* Replaces MetalStereo::compute_raw (metal_stereo.hpp:79-81, metal_stereo.cpp:340-349): rectifies raw
 * host frames with the resident maps -- calib::remap (libs/calib/src/rectify.cpp:133-148) at the
 * maps' resolution and downsample2 (libs/depth/src/stereo.cpp:348-359), fused in one kernel and
 * bit-exact with the two -- then runs sgm_compute's pipeline on the result. rect_left / rect_right
 * (may be NULL) receive the rectified pair, (map_w / 2) x (map_h / 2) with pitch map_w / 2.
 * Errors as the reference: "set_rectification() not called", "raw size mismatch". */
int sgm_compute_raw(sgm_context* ctx, const uint8_t* raw_left, const uint8_t* raw_right, ptrdiff_t raw_stride, int raw_w,
                    int raw_h, const sgm_outputs* out, uint8_t* rect_left, uint8_t* rect_right);

/* The same on device-resident raw frames and outputs, enqueued on `stream` (NULL = the context's)
 * without waiting. d_rect_left / d_rect_right (pitch map_w / 2) receive the rectified pair; NULL
 * uses the context's input buffers. */
int sgm_compute_raw_device(sgm_context* ctx, const uint8_t* d_raw_left, const uint8_t* d_raw_right, ptrdiff_t raw_pitch,
                           int raw_w, int raw_h, const sgm_outputs* d_out, uint8_t* d_rect_left, uint8_t* d_rect_right,
                           void* stream);

/* depth::StereoParams (libs/depth/include/einstar/depth/stereo.hpp:42-89): SGM at 1/2^pyramid_levels
 * of the input, ZNCC refinement to each finer level, detail refinement on the full-resolution pair,
 * occlusion and spike filters, speckle removal. sgm.subpixel is StereoParams::subpixel. */
typedef struct {
    sgm_params sgm;
    int32_t pyramid_levels;              /* 0..4 */
    int32_t refine_search_radius;        /* RefineParams (stereo.hpp:42-48): 1..4 */
    int32_t refine_zncc_radius;          /* 1..7 */
    float refine_min_zncc;
    int32_t refine_reject_edge_winners;
    int32_t detail_enabled;              /* DetailParams (stereo.hpp:55-64) */
    int32_t detail_radius;               /* 1..16 */
    float detail_search, detail_step;    /* lround(search / step) <= 8 */
    float detail_min_zncc;
    int32_t filter_occlusion;            /* FilterParams (stereo.hpp:69-74) */
    float filter_occlusion_tolerance;
    int32_t filter_spikes;
    float filter_spike_threshold;
    int32_t speckle_max_region_size;     /* SpeckleParams (stereo.hpp:76-79) */
    float speckle_max_diff;
} sgm_stereo_params;

/* Fills the reference defaults (stereo.hpp:26-89: two pyramid levels, every stage on). */
void sgm_stereo_params_default(sgm_stereo_params* p);

/* Replaces MetalStereo::create with its full StereoParams (metal_stereo.hpp:70-71, checks
 * metal_stereo.cpp:115-149): a context for depth::compute_disparity on width x height images (the
 * finest level); SGM runs at (width >> levels) x (height >> levels). Everything is allocated here. */
int sgm_create_stereo(const sgm_stereo_params* params, int width, int height, sgm_context** out);

/* Replaces MetalStereo::compute (metal_stereo.hpp:77-78) = depth::compute_disparity
 * (libs/depth/src/stereo.cpp:496-536), both forms: full_left / full_right (may be NULL) are the
 * full_width x full_height images left and right were box-downsampled from, used by the detail
 * refinement when full_width == 2 * width and full_height == 2 * height (otherwise skipped, as in
 * the reference). Host images in, host disparity and confidence (may be NULL) out; blocking. On a
 * context from sgm_create it is sgm_disparity with confidence 1 where valid. */
int sgm_compute_disparity(sgm_context* ctx, const uint8_t* left, const uint8_t* right, ptrdiff_t stride, int width,
                          int height, const uint8_t* full_left, const uint8_t* full_right, ptrdiff_t full_stride,
                          int full_width, int full_height, float* disparity, float* confidence);

/* The same on device-resident images and outputs, enqueued on `stream` without waiting. */
int sgm_compute_disparity_device(sgm_context* ctx, const uint8_t* d_left, const uint8_t* d_right, ptrdiff_t pitch,
                                 int width, int height, const uint8_t* d_full_left, const uint8_t* d_full_right,
                                 ptrdiff_t full_pitch, int full_width, int full_height, float* d_disparity,
                                 float* d_confidence, void* stream);

/* Replaces MetalStereo::compute_raw with its full StereoParams (metal_stereo.cpp:340-349, 448-479):
 * raw frames rectified with the resident maps (sgm_set_rectification; maps at twice the context's
 * width x height) to the full-resolution pair when the detail refinement is on (then
 * downsample2 of it) or straight to the stereo resolution, then compute_disparity. rect_left /
 * rect_right (may be NULL) receive the stereo-resolution rectified pair. */
int sgm_compute_disparity_raw(sgm_context* ctx, const uint8_t* raw_left, const uint8_t* raw_right, ptrdiff_t raw_stride,
                              int raw_w, int raw_h, float* disparity, float* confidence, uint8_t* rect_left,
                              uint8_t* rect_right);

/* The same on device-resident raw frames and outputs, enqueued on `stream` without waiting. */
int sgm_compute_disparity_raw_device(sgm_context* ctx, const uint8_t* d_raw_left, const uint8_t* d_raw_right,
                                     ptrdiff_t raw_pitch, int raw_w, int raw_h, float* d_disparity, float* d_confidence,
                                     uint8_t* d_rect_left, uint8_t* d_rect_right, void* stream);

/* depth::RectifiedGeometry (libs/depth/include/einstar/depth/point_image.hpp:12-26) at the disparity's
 * resolution: focal length and left principal point in pixels, baseline in mm, right cx - left cx. */
typedef struct {
    double f, cx, cy, baseline, cx_offset;
} sgm_rectified_geometry;

/* depth::PointImageParams (point_image.hpp:43-49). */
typedef struct {
    float min_depth, max_depth;   /* mm */
    float max_depth_jump;         /* mm between the neighbours used for a normal */
    float edge_weight;            /* view weight of a point without a normal */
    float weight_reference_depth; /* mm; 0 = no depth-noise term */
} sgm_point_params;

/* Fills the reference defaults (point_image.hpp:43-49). */
void sgm_point_params_default(sgm_point_params* p);

/* An organised point frame, w*h each: points and normals as (x, y, z, valid) float4s in the
 * rectified left camera frame (zero where invalid; a normal's fourth value is 1 where it exists),
 * weights as floats. As MetalFrameData holds them (points_xyzw, normals_xyzw, weights). */
typedef struct {
    float* points_xyzw;
    float* normals_xyzw;
    float* weights;
} sgm_frame;

/* Replaces depth::disparity_to_points (point_image.hpp:57-58, point_image.cpp:10-55) on device data:
 * d_confidence may be NULL (weight 1). Enqueued on `stream` (NULL = the null stream), no wait;
 * allocates nothing. Points are formed in double as the CPU forms them. */
int sgm_disparity_to_points_device(const float* d_disparity, const float* d_confidence, int width, int height,
                                   const sgm_rectified_geometry* geometry, const sgm_point_params* params,
                                   const sgm_frame* d_out, void* stream);

/* Replaces depth_metal::DepthPacker::pack (libs/depth_metal/include/einstar/depth_metal/depth_packer.hpp:17-38,
 * depth_packer.cpp:12-26): a device point frame into the session file's uncompressed depth layout
 * (libs/session/src/session.cpp:143-160): w*h u16 row-delta codes of lround(max(0, z) * 50) clamped
 * to [1, 65535] (0 where z <= 0), then w*h u8 lround(weight * 255) clamped to [0, 255]; 3*w*h bytes
 * at d_out. Quantised in double as the session writer does (the Metal packer uses float and may
 * differ by one unit). d_weights may be NULL (confidence 255). No wait; allocates nothing. */
int sgm_pack_depth_device(const float* d_points_xyzw, const float* d_weights, int width, int height, uint8_t* d_out,
                          void* stream);

/* Device workspace sgm_remove_speckles_device needs for a width x height map: 12 bytes a pixel (labels,
 * tile-local counts, region sizes); 0 for an empty size. */
size_t sgm_speckle_workspace_bytes(int width, int height);

/* Replaces depth::remove_speckles (stereo.hpp:117-118, stereo.cpp:465-494) on device data: d_out receives
 * d_disparity with every region (4-neighbourhood, valid neighbours |d_a - d_b| <= max_diff) of fewer
 * than max_region_size pixels set to the invalid disparity; max_region_size <= 1 copies. d_out may be
 * d_disparity. d_work: sgm_speckle_workspace_bytes(width, height) bytes, 4-byte aligned (unused and
 * may be NULL when max_region_size <= 1). Enqueued on `stream` (NULL = the null stream), no wait;
 * allocates nothing. At most 2^31 - 1 pixels. */
int sgm_remove_speckles_device(const float* d_disparity, float* d_out, int width, int height, int max_region_size,
                               float max_diff, void* d_work, void* stream);

/* Replaces depth::filter_disparity (stereo.hpp:124-125, stereo.cpp:286-326) on device data, in place:
 * the occlusion filter (FilterParams::occlusion, occlusion_tolerance), then the spike filter (spikes,
 * spike_threshold), as the reference orders them. d_work: width*height floats for the spike filter's
 * output, which is then copied back (may be NULL when spikes == 0). Widths up to 32768 (rows beyond 4096
 * columns need 4 bytes of LDS a column for the occlusion filter). Enqueued on `stream`, no wait;
 * allocates nothing. */
int sgm_filter_disparity_device(float* d_disparity, int width, int height, int occlusion, float occlusion_tolerance,
                                int spikes, float spike_threshold, float* d_work, void* stream);

/* Replaces MetalStereo::set_point_params (metal_stereo.hpp:83-85): the geometry and limits the frame
 * entry points convert disparity with; allocates the context's point buffers. */
int sgm_set_point_params(sgm_context* ctx, const sgm_rectified_geometry* geometry, const sgm_point_params* params);

/* Replaces MetalStereo::compute_frame_raw (metal_stereo.hpp:86-90, metal_stereo.cpp:250-279): raw
 * frames -> rectified pair -> compute_disparity -> points, normals, weights. Host outputs (any of the
 * three may be NULL) and rectified pair (may be NULL); blocking. Errors as the reference:
 * "set_point_params() not called", "set_rectification() not called", "raw size mismatch". */
int sgm_compute_frame_raw(sgm_context* ctx, const uint8_t* raw_left, const uint8_t* raw_right, ptrdiff_t raw_stride,
                          int raw_w, int raw_h, const sgm_frame* out, uint8_t* rect_left, uint8_t* rect_right);

/* The same on device-resident raw frames and outputs (NULL outputs: the context's buffers), on
 * `stream` without waiting. */
int sgm_compute_frame_raw_device(sgm_context* ctx, const uint8_t* d_raw_left, const uint8_t* d_raw_right,
                                 ptrdiff_t raw_pitch, int raw_w, int raw_h, const sgm_frame* d_out, uint8_t* d_rect_left,
                                 uint8_t* d_rect_right, void* stream);

/* The context's own stream (a hipStream_t), on which calls with stream == NULL are enqueued, so that a
 * caller can order its own work (e.g. sgm_pack_depth_device) after them. No counterpart in the
 * reference (its command buffers are internal). */
void* sgm_get_stream(sgm_context* ctx);

/* Waits for the work enqueued by this context and records its timings. */
int sgm_synchronize(sgm_context* ctx);

/* Replaces MetalStereo::last_timings (metal_stereo.hpp:97). */
int sgm_last_timings(const sgm_context* ctx, sgm_timings* out);

/* Stage times summed over the calls completed since the last reset (read at sgm_synchronize or
 * sgm_compute). No counterpart in the reference (its profiler serialises the GPU per stage,
 * libs/gpu/include/einstar/gpu/profile.hpp:1-6). */
int sgm_get_stage_totals(const sgm_context* ctx, sgm_stage_totals* out);
int sgm_reset_stage_totals(sgm_context* ctx);

/* Tuning knobs: they change speed and memory, never results. */
enum {
    SGM_OPT_COST_VOLUME = 1, /* 1 (default): matching costs written once as bytes and read by every
                                path (W*H*dp bytes of HBM); 0: recomputed from census words in each path */
    SGM_OPT_PATH_LINES = 2,  /* scanlines per path-kernel block: 4, 8 or 16, or 0 (default): with 16
                                lanes per scanline 16, or 4 for 4-path pairs of at most 640x400 pixels,
                                more than 16 disparities a lane or costs mostly recomputed; with 32 lanes
                                8, or 4 for 4-path pairs of 384 or more padded disparities; with 64, 4 */
    SGM_OPT_PATH_GROUP = 3,  /* lanes per scanline in the path kernel: 0 (default) = chosen per pair
                                (64 where the padded disparity count is a multiple of 256 and above 384,
                                or a multiple of 256 with half the directions or more recomputing their
                                costs; else 32 where it is a multiple of 128 and the pair recomputes half
                                its directions or more, has 4 paths and at most 1M pixels, or has 8
                                paths, 256 or more padded disparities and at most 640x512 pixels; else
                                16), or 16, or 32 (used when the lane layout allows it -- padded
                                disparities a multiple of 64 -- with at most 8 lines per block, else 16),
                                or 64 (used with 4 lines per block where the padded disparities are a
                                multiple of 256, else as 32) */
    SGM_OPT_RECOMPUTE_DIRS = 4, /* with the cost volume on: how many of the path directions (in the order
                                  left-right, right-left, down, up, then the diagonals) recompute their
                                  costs from census words instead of reading them, 0..8 */
    SGM_OPT_VIEW_BASE_HIGH = 5, /* test aid, value 1: re-places the census words, step penalties and
                                  cost volume (the buffers the path kernel reads through 32-bit buffer
                                  views) at addresses whose low 32-bit word has bit 31 set, each inside
                                  an allocation 4 GiB larger than it, so that a sign extension of that
                                  word would show up as wrong results */
    SGM_OPT_ROW_THREADS = 6  /* threads per row-kernel block (one image row a block): 0 (default) = 1024
                                for pairs of at most 256 rows, else 256; or 256, or 1024 (512 with 8
                                paths and 256 or more padded disparities, and with 512; rows of more
                                than 4096 columns and u16 path costs always run 256) */
};
int sgm_set_option(sgm_context* ctx, int option, int64_t value);

/* Device memory the context holds, in bytes. */
size_t sgm_device_bytes(const sgm_context* ctx);

/* Device memory helpers for callers without a HIP binding of their own (ctypes, cgo, JNI):
 * thin wrappers of hipMalloc / hipFree / hipMemcpy on the current device. No counterpart in the
 * reference (Metal buffers come from gpu::Context, libs/gpu/include/einstar/gpu/context.hpp:70-74). */
int sgm_dev_malloc(size_t bytes, void** out);
int sgm_dev_free(void* ptr);
int sgm_dev_upload(void* dst, const void* src, size_t bytes);
int sgm_dev_download(void* dst, const void* src, size_t bytes);
int sgm_dev_memset(void* dst, int value, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte aligned pointers and size); also the
 * known-size reference for HBM traffic counters. Blocks until done. */
int sgm_dev_copy(void* dst, const void* src, size_t bytes);
/* Pinned (page-locked) host memory: device copies to and from it are direct DMA transfers. */
int sgm_host_alloc(size_t bytes, void** out);
int sgm_host_free(void* ptr);
/* Free and total device memory in bytes (hipMemGetInfo). */
int sgm_dev_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ------------------------------------------------------------------------------------------
 * Batched SGM: many rectified pairs of uneven sizes and disparity ranges in one graph launch.
 * No counterpart in the reference (its frontend runs one pair per call,
 * libs/pipeline/src/stereo_frontend.cpp:302-348); every pair's result equals sgm_compute's.
 * ------------------------------------------------------------------------------------------ */

typedef struct {
    const uint8_t* left;     /* device */
    const uint8_t* right;    /* device */
    int64_t pitch;           /* bytes */
    int32_t width, height;
    int32_t min_disparity;   /* per-pair disparity range; other parameters come from the batch */
    int32_t num_disparities;
    sgm_outputs out;         /* device pointers; disparity is required */
} sgm_pair;

typedef struct sgm_batch sgm_batch;

/* Launch plan knobs: -1 = let the autotuner choose. None changes any result. */
typedef struct {
    int32_t cost_volume; /* 0: costs recomputed in the path kernel; 1: written once as bytes */
    int32_t path_lines;  /* scanlines per path-kernel block: 4, 8 or 16 */
    int32_t branches;    /* 0: one stream; 1: a graph branch per lane class; 2: two per class */
    int32_t path_group;  /* lanes per scanline: 16, or 32 (with at most 8 lines; lane classes that do not
                            allow it use 16) */
    int32_t sweep_band;  /* 0: every direction walked line by line; > 0: band sweeps (8 paths, byte path
                            costs, cost volume): the six non-horizontal directions aggregated row by row,
                            this many rows per launch, into one u16 sum per sweep */
    int32_t sweep_tile;  /* band sweeps: columns per block, 16, 32, 48 or 64 (lane classes whose LDS
                            would not hold it use less) */
} sgm_batch_tuning;

/* The plan a batch runs and how it was chosen. */
typedef struct {
    int32_t cost_volume, path_lines, branches;
    int32_t candidates_timed; /* plans timed at creation (0 when taken from the cache) */
    double best_ms, worst_ms; /* replay time of the fastest and slowest candidate */
    int32_t from_cache;
    int32_t path_group;       /* lanes per scanline (16 or 32) */
    int32_t sweep_band;       /* rows per band-sweep launch; 0: line schedule */
    int32_t sweep_tile;       /* band sweeps: columns per block (unused by the line schedule) */
} sgm_batch_plan;

/* Plans a batch of n pairs: allocates the scratch of all pairs, times every candidate plan on the
 * batch itself (a graph replay each) and keeps the fastest, or takes the plan cached for its key
 * (sgm_tune_cache_load below); the chosen launches are captured in a hipGraph. The pair images and
 * outputs must stay valid until sgm_batch_destroy; their contents may change between runs. */
int sgm_batch_create(const sgm_params* params, int n, const sgm_pair* pairs, sgm_batch** out);
int sgm_batch_create_tuned(const sgm_params* params, int n, const sgm_pair* pairs, const sgm_batch_tuning* tuning,
                           sgm_batch** out);
int sgm_batch_get_plan(const sgm_batch* b, sgm_batch_plan* out);
/* The autotuner's plan cache. sgm_batch_create caches the plan of every search that left all knobs
 * open, keyed by the parameters that select kernel variants (path count, census window, slant steps)
 * and the pairs' shapes bucketed to 64 pixels with their lane layout; a later batch of the same key
 * takes the cached plan (from_cache = 1) without timing candidates, where it honours the knobs that
 * batch fixes. load merges a file written by save into the process's cache (*entries = lines read);
 * save writes the whole cache (*entries = lines written); entries may be NULL. No counterpart in the
 * reference. */
int sgm_tune_cache_load(const char* path, int* entries);
int sgm_tune_cache_save(const char* path, int* entries);
/* Replays the graph on `stream` (hipStream_t; NULL = the batch's own stream) without waiting. */
int sgm_batch_run(sgm_batch* b, void* stream);
/* The same work launched pair by pair without batching or graphs: the "plain" reference point. */
int sgm_batch_run_plain(sgm_batch* b, void* stream);
int sgm_batch_synchronize(sgm_batch* b);
const char* sgm_batch_last_error(const sgm_batch* b);
size_t sgm_batch_device_bytes(const sgm_batch* b);
void sgm_batch_destroy(sgm_batch* b);

/* ------------------------------------------------------------------------------------------
 * Serving: a request loop with continuous admission. No counterpart in the reference (its live
 * pipeline queues frames for one worker, libs/pipeline/src/scan_pipeline.cpp:512-609).
 *
 * sgm_server_start reserves an HBM arena (hbm_budget_bytes; 0 = half the free HBM, at most 64 GiB)
 * and `slots` streams with worker threads. A submitted request (host images, copied at submission
 * unless it borrows them) waits in arrival order; the head is admitted when a slot is free and
 * its device footprint fits the arena. Frames arriving while others run start as soon as a slot and
 * memory free up and overlap on the device. Results are copied straight into the request's host
 * output buffers (pinned ones, e.g. from sgm_host_alloc, take a direct DMA), which must stay valid
 * until its completion is polled. All functions are thread-safe.
 * ------------------------------------------------------------------------------------------ */

typedef struct {
    int32_t max_width, max_height; /* largest request */
    int32_t max_disparities;       /* largest num_disparities of a request */
    int32_t slots;                 /* requests on the device at once (1..64) */
    int64_t hbm_budget_bytes;      /* arena size; 0 = default */
    int32_t compute_streams;       /* 0 (default) = 1: requests' kernels run one request at a time, in
                                      admission order, while other requests' copies overlap them;
                                      k in 1..slots: k shared compute streams; -1: each slot computes on
                                      its own stream (requests' kernels share the GPU) */
    int32_t path_group;            /* lanes per path scanline: 0 (default) = chosen per request shape by
                                      timing 16, 32 and 64 lanes (those the shape's lane layout allows) on
                                      the first requests of that shape; 16, 32 or 64 = fixed (as
                                      SGM_OPT_PATH_GROUP). Results never depend on it. */
} sgm_server_config;

typedef struct {
    const uint8_t* left;  /* host */
    const uint8_t* right; /* host */
    int64_t stride;       /* bytes */
    int32_t width, height;
    int32_t min_disparity, num_disparities;
    sgm_outputs out; /* host buffers; disparity required; no volumes or census */
    uint64_t user_tag;
    int32_t rectification; /* 0: left/right are rectified width x height images. k > 0: they are raw
                              frames of map set k (sgm_server_add_rectification), rectified on the
                              device; width x height must then be half the maps' size */
    int32_t borrow_inputs; /* 0: left/right are copied at submission and may be reused at once.
                              1: they are read in place (a direct DMA when pinned, e.g. from
                              sgm_host_alloc) and must stay valid and unchanged until the
                              completion is polled, as MetalStereo::compute_frame uses page-aligned
                              raw images in place (metal_stereo.hpp:90-93) */
} sgm_request;

typedef struct {
    uint64_t ticket;
    uint64_t user_tag;
    int32_t status;      /* SGM_OK or the request's error */
    int64_t t_submit_ns; /* sgm_server_now_ns() clock */
    int64_t t_admit_ns;
    int64_t t_done_ns;   /* results in the caller's buffers */
    double gpu_ms;       /* first to last kernel of this request */
} sgm_completion;

typedef struct {
    int32_t queued;
    int32_t in_flight;
    int64_t completed;
    int64_t failed;
    int64_t arena_bytes;
    int64_t arena_free_bytes;
} sgm_server_stats;

typedef struct sgm_server sgm_server;

int sgm_server_start(const sgm_params* params, const sgm_server_config* cfg, sgm_server** out);
/* SGM_ERR_BUSY when the request could never fit the arena; SGM_ERR_INVALID_ARGUMENT for sizes,
 * ranges or outputs the server does not take. */
int sgm_server_submit(sgm_server* sv, const sgm_request* rq, uint64_t* ticket);
/* Registers a rectification map set (both cameras' calib::RemapTable, as sgm_set_rectification)
 * for raw frames of one size and returns its id (> 0) in *id. The maps stay resident on the device
 * until sgm_server_stop, as MetalStereo keeps those of set_rectification (metal_stereo.cpp:203-231)
 * across frames; requests name the set in sgm_request.rectification. */
int sgm_server_add_rectification(sgm_server* sv, const float* left_map_x, const float* left_map_y,
                                 const float* right_map_x, const float* right_map_y, int map_w, int map_h, int raw_w,
                                 int raw_h, int32_t* id);
/* Up to max completions; waits up to timeout_us (0: no wait, < 0: until one is available). */
int sgm_server_poll(sgm_server* sv, sgm_completion* out, int max, int timeout_us, int* n);
int sgm_server_get_stats(sgm_server* sv, sgm_server_stats* st);
const char* sgm_server_last_error(sgm_server* sv);
/* Finishes the requests on the device, drops the queued ones, frees everything. */
int sgm_server_stop(sgm_server* sv);
/* The steady clock (ns) the completion timestamps use. */
int64_t sgm_server_now_ns(void);

#ifdef __cplusplus
}
#endif

#endif /* SGM_HIP_H */
