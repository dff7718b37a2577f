/*
 * CPU oracle for the SGM stereo hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference CPU matcher's semi-global matching subset
 * (libs/depth/src/stereo.cpp:15-96, 129-147, 328-346, 361-463 and
 * libs/depth/include/einstar/depth/stereo.hpp:22-40, 107-115). It is the checker the
 * HIP library is compared against and the timed CPU baseline; nothing in the shipped
 * library links, loads or calls it. Only tests, the smoke check and the CPU-baseline
 * leg of bench.py may use it.
 *
 * Every function names the reference lines it follows. Integer semantics are kept
 * exactly (u16 costs and sums, first-strict-argmin ties, the `break` in the cost fill,
 * lround in the left/right check, nth_element's n/2-th value in median3). Float code is
 * compiled with -ffp-contract=off so it rounds exactly like the reference.
 *
 * Parity pins (tests/test_oracle_*.py): census bit order by hand-built windows; cost by
 * popcount closed forms and image-bound cases; aggregation by exhaustive enumeration of
 * every disparity sequence along each path on tiny images, the P1=P2=0 closed form and
 * per-direction coverage; WTA/uniqueness/subpixel by crafted sum vectors and the
 * parabola vertex; left/right check by crafted maps including exact .5 ties; median3 by
 * brute-force sorting; the whole path by the reference's own fronto-parallel quality bar
 * (tests/depth/test_stereo.cpp:121-141). The reference CPU path itself cannot be built
 * here (it includes <tbb/parallel_for.h>, whose headers the image lacks).
 */
#ifndef SGM_ORACLE_H
#define SGM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors depth::SgmParams (stereo.hpp:26-40) plus StereoParams::subpixel (stereo.hpp:87). */
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
} orc_params;

/* kInvalidDisparity (stereo.hpp:23) and the right-WTA "no match" value (stereo.cpp:413). */
#define ORC_INVALID_DISPARITY (-1.0e9f)
#define ORC_NO_MATCH INT32_MIN

/* Per-pixel status bits of the left WTA chain (stereo.cpp:419-461). */
#define ORC_ST_WINDOW 1u  /* candidate window [min_k, max_k) non-empty */
#define ORC_ST_UNIQUE 2u  /* passed the uniqueness test */
#define ORC_ST_LR 4u      /* passed the left/right check */
#define ORC_ST_VALID 7u

typedef struct {
    float* disparity;         /* W*H: sgm_disparity() output, ORC_INVALID_DISPARITY where invalid */
    float* disparity_median;  /* W*H: median3(disparity); may be NULL */
    int32_t* disparity_int;   /* W*H: min_d + best_k, ORC_NO_MATCH for an empty window; may be NULL */
    uint8_t* status;          /* W*H: ORC_ST_* bits; may be NULL */
    int32_t* right_disparity; /* W*H: right WTA, ORC_NO_MATCH where none; may be NULL */
    uint16_t* best_cost;      /* W*H: S at best_k, 0xFFFF for an empty window; may be NULL */
    uint16_t* second_cost;    /* W*H: best non-adjacent S, 0xFFFF if none; may be NULL */
    uint16_t* sum_volume;     /* W*H*D: aggregated S [y][x][d]; may be NULL */
    uint16_t* cost_volume;    /* W*H*D: matching cost C [y][x][d]; may be NULL */
    uint64_t* census_left;    /* W*H; may be NULL */
    uint64_t* census_right;   /* W*H; may be NULL */
} orc_outputs;

/* threads <= 0 means one per online CPU. Results never depend on the thread count. */

/* census_transform, stereo.cpp:328-346. */
void orc_census(const uint8_t* img, int w, int h, ptrdiff_t stride, int rx, int ry, uint64_t* out, int threads);

/* Cost fill inside sgm_disparity, stereo.cpp:366-380. */
void orc_cost_volume(const uint64_t* cl, const uint64_t* cr, int w, int h, int min_d, int nd, uint16_t invalid_cost,
                     uint16_t* cost, int threads);

/* One direction of aggregate_path over every start pixel of that direction, adding into sum
 * (stereo.cpp:44-89 and the start set of :390-408). */
void orc_aggregate_direction(const uint16_t* cost, const uint8_t* guide, ptrdiff_t guide_stride, int w, int h, int nd,
                             const orc_params* p, int dx, int dy, uint16_t* sum, int threads);

/* All 4 or 8 directions into a zeroed sum volume (stereo.cpp:385-409). */
void orc_aggregate(const uint16_t* cost, const uint8_t* guide, ptrdiff_t guide_stride, int w, int h, const orc_params* p,
                   uint16_t* sum, int threads);

/* subpixel_offset, stereo.cpp:92-96. */
float orc_subpixel_offset(float cm, float c0, float cp);

/* Left WTA with uniqueness and subpixel, and the right WTA, stereo.cpp:411-448. */
void orc_wta(const uint16_t* sum, int w, int h, const orc_params* p, float* disp, int32_t* right_disp,
             int32_t* disp_int, uint8_t* status, uint16_t* best_cost, uint16_t* second_cost, int threads);

/* Left/right consistency, stereo.cpp:450-461 (in place on disp; status gains ORC_ST_LR). */
void orc_lr_check(float* disp, const int32_t* right_disp, int w, int h, int lr_max_diff, uint8_t* status, int threads);

/* median3, stereo.cpp:129-147. */
void orc_median3(const float* in, int w, int h, float* out, int threads);

/* calib::remap, libs/calib/src/rectify.cpp:133-148: bilinear lookup of src at (map_x, map_y) per
 * output pixel, 0 where the 2x2 neighbourhood leaves src, rounded by +0.5 and truncation. */
void orc_remap(const uint8_t* src, int src_w, int src_h, ptrdiff_t src_stride, const float* map_x, const float* map_y,
               int w, int h, uint8_t* dst, int threads);

/* downsample2, stereo.cpp:348-359: (a + b + c + d + 2) / 4 over 2x2 blocks; output (w/2) x (h/2). */
void orc_downsample2(const uint8_t* src, int w, int h, ptrdiff_t stride, uint8_t* dst, int threads);

/* The stereo front end's rectification to the SGM level (libs/pipeline/src/stereo_frontend.cpp:331-334):
 * remap at the maps' resolution (map_w x map_h), then downsample2. dst is (map_w/2) x (map_h/2). */
int orc_rectify_half(const uint8_t* raw, int raw_w, int raw_h, ptrdiff_t raw_stride, const float* map_x, const float* map_y,
                     int map_w, int map_h, uint8_t* dst, int threads);

/* depth::StereoParams (stereo.hpp:42-89) minus its SgmParams, which travel as orc_params (whose
 * subpixel field is StereoParams::subpixel and applies to every stage). */
typedef struct {
    int32_t pyramid_levels;
    int32_t refine_search_radius, refine_zncc_radius; /* RefineParams (stereo.hpp:42-48) */
    float refine_min_zncc;
    int32_t refine_reject_edge_winners;
    int32_t detail_enabled, detail_radius;             /* DetailParams (stereo.hpp:55-64) */
    float detail_search, detail_step, detail_min_zncc;
    int32_t filter_occlusion;                          /* FilterParams (stereo.hpp:69-74) */
    float filter_occlusion_tolerance;
    int32_t filter_spikes;
    float filter_spike_threshold;
    int32_t speckle_max_region_size;                   /* SpeckleParams (stereo.hpp:76-79) */
    float speckle_max_diff;
} orc_stereo_params;

/* zncc_slanted, stereo.cpp:101-125 (-2: window leaves b, -1: a flat window). */
float orc_zncc_slanted(const uint8_t* a, const uint8_t* b, ptrdiff_t stride, int bw, int x, int y, float d, float gx,
                       float gy, int r);

/* refine_level, stereo.cpp:151-207: out (w x h) from coarse_raw (cw x ch, median3 applied inside). */
void orc_refine_level(const uint8_t* left, const uint8_t* right, ptrdiff_t stride, int w, int h, const float* coarse_raw,
                      int cw, int ch, const orc_stereo_params* sp, int subpixel, float* out_disp, float* out_conf,
                      int threads);

/* zncc_window, stereo.cpp:211-239. */
float orc_zncc_window(const uint8_t* a, const uint8_t* b, ptrdiff_t stride, int bw, int x0, int y0, int n, float cx,
                      float cy, float D, float gx, float gy);

/* refine_detail, stereo.cpp:243-284, in place on disparity (w x h) with full images (W x H). */
void orc_refine_detail(const uint8_t* full_left, const uint8_t* full_right, ptrdiff_t full_stride, int W, int H,
                       float* disparity, int w, int h, const orc_stereo_params* sp, int subpixel, int threads);

/* filter_disparity, stereo.cpp:286-326 (occlusion then spikes), in place. */
void orc_filter_disparity(float* disparity, int w, int h, const orc_stereo_params* sp, int threads);

/* remove_speckles, stereo.cpp:465-494, in place. Returns 0, or -1 when an allocation fails. */
int orc_remove_speckles(float* disparity, int w, int h, const orc_stereo_params* sp);

/* compute_disparity, stereo.cpp:496-536 (both forms: full_left/full_right may be NULL). disparity and
 * confidence are w x h. Returns 0, or -1 when an allocation fails. */
int orc_compute_disparity(const uint8_t* left, const uint8_t* right, ptrdiff_t stride, int w, int h,
                          const uint8_t* full_left, const uint8_t* full_right, ptrdiff_t full_stride, int full_w,
                          int full_h, const orc_params* sgm, const orc_stereo_params* sp, float* disparity,
                          float* confidence, int threads);

/* depth::RectifiedGeometry (libs/depth/include/einstar/depth/point_image.hpp:12-26). */
typedef struct {
    double f, cx, cy, baseline, cx_offset;
} orc_geometry;

/* depth::PointImageParams (point_image.hpp:43-49). */
typedef struct {
    float min_depth, max_depth, max_depth_jump, edge_weight, weight_reference_depth;
} orc_point_params;

/* disparity_to_points, libs/depth/src/point_image.cpp:10-55: points and normals as [y][x][3] floats
 * (zero where invalid), weights [y][x]. confidence may be NULL (weight 1). */
void orc_disparity_to_points(const float* disparity, const float* confidence, int w, int h, const orc_geometry* g,
                             const orc_point_params* p, float* points, float* normals, float* weights, int threads);

/* The session file's uncompressed depth layout (libs/session/src/session.cpp:143-160, with
 * capture_depth's depth = max(0, z) and confidence = weight, :296-313): w*h u16 row-delta codes of
 * lround(z * 50) clamped to [1, 65535] (0 where z <= 0), then w*h u8 lround(weight * 255) clamped to
 * [0, 255]. z is read from points [y][x][3] (index 2); weights may be NULL (then 1). out: 3*w*h bytes. */
void orc_pack_depth(const float* points, const float* weights, int w, int h, uint8_t* out);

/* decode_images without the zstd step (session.cpp:164-185): depth in mm and confidence. */
void orc_unpack_depth(const uint8_t* in, int w, int h, float* depth, float* confidence);

/* sgm_disparity (stereo.cpp:361-463) followed by median3; fills whatever outputs are non-NULL.
 * Returns 0, or -1 when an allocation fails. */
int orc_sgm(const uint8_t* left, const uint8_t* right, ptrdiff_t stride, int w, int h, const orc_params* p,
            const orc_outputs* out, int threads);

#ifdef __cplusplus
}
#endif

#endif
