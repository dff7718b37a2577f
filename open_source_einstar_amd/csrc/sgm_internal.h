// Internal types shared by the gfx950 SGM kernels (sgm_kernels.hip) and the C-ABI layer
// (sgm_api.hip). Not installed; the public interface is include/sgm_hip.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgm {

// Lanes that cooperate on one scanline: one 16-lane DPP row of a wave64. Each lane owns `dpl`
// consecutive disparities, so a line covers dp = 16 * dpl >= D disparities.
constexpr int kGroup = 16;
constexpr int kBlock = 256;                   // threads per block (4 wave64s)
constexpr int kLinesPerBlock = kBlock / kGroup;
constexpr int kMaxBatch = 64;                 // pairs per launch table
constexpr int kMaxDirs = 8;
// Widest image row: the row and occlusion kernels keep 4 bytes of LDS per column (128 KiB here);
// rows up to kMaxStagedWidth also stage their left results there (15 bytes per column).
constexpr int kMaxWidth = 32768;
constexpr int kMaxStagedWidth = 4096;
// The path kernel reads the census rows, the step penalties and the cost volume through buffer
// views addressed by 32-bit byte offsets, so each of them must stay within 4 GiB - 1. Census rows
// beyond that are refused (the penalties are half their size); a cost volume beyond it is not used
// (costs are recomputed).
constexpr uint64_t kMaxViewBytes = 0xFFFFFFFFull;
constexpr int kMaxLevels = 4;                 // pyramid levels below the full resolution

// Device pointers and geometry of one rectified pair. Every output pointer may be null except
// disp, which the left/right check and median read back.
struct PairDesc {
    const uint8_t* left;
    const uint8_t* right;
    uint64_t* census_l;
    uint64_t* census_r;
    uint32_t* edge;          // [y][x] left-image step penalties of the four path axes, one byte each
    uint8_t* lvol;           // [y][x][dir][dp] per-direction path costs L (u8)
    uint8_t* cvol;           // [y][x][dp] matching costs C (u8), or null: the path kernel recomputes them
    float* disp;             // sgm_disparity output (post left/right check)
    float* median;           // median3(disp)
    float* conf;             // 1 where disp is valid, else 0
    int32_t* dint;           // min_d + best_k, INT32_MIN for an empty candidate window
    uint8_t* status;         // SGM_STATUS_* bits
    int32_t* rdisp;          // right WTA, INT32_MIN for no match
    uint16_t* best;          // S at best_k
    uint16_t* second;        // best non-adjacent S (0xFFFF: none)
    uint16_t* sum_volume;    // optional debug S [y][x][nd]
    uint16_t* cost_volume;   // optional debug C [y][x][nd]
    uint8_t* seed;           // band sweeps: [sweep][parity][kind][x][dp] path-cost bytes of a band's last row
    int64_t pitch;           // input row pitch in bytes
    int32_t census_pitch;    // census words per row: pads, then the w words of the row, then pads
    int32_t census_padl;     // words before x = 0 (every disparity window stays inside the row)
    int64_t lvol_px_stride;  // bytes per pixel of lvol: ndirs * dp
    int32_t w, h;
    int32_t min_d, nd, dpl, dp;
    int32_t ndirs;
    int32_t total_lines;             // scanlines over all directions
    int32_t line_off[kMaxDirs];      // first line index of each direction
};

// Parameters shared by every pair of a launch (depth::SgmParams minus the disparity range).
struct SgmConsts {
    int32_t p1, p2, adaptive_p2, slant_steps;
    int32_t wide_costs;   // 1: path costs and step penalties are u16 (see wide_path_costs), 0: bytes
    int32_t ndirs;  // 4 or 8 path directions
    int32_t cv_from_dir;  // cost-volume launches: directions below this recompute their costs
    int32_t census_rx, census_ry, invalid_cost;
    int32_t lr_max_diff, subpixel;
    float uniqueness;
    int32_t p2_by_q[32];  // max(P1+1, P2/(q+1)) for q = |dI|/8 in [0, 31]
};

// Maps a flat block index of one launch to (pair, block within the pair). A single pair travels
// inline in the kernel arguments; batches point at a device-resident descriptor array.
struct LaunchTable {
    PairDesc one;                 // used when n == 1
    const PairDesc* many;         // device array of n descriptors (n > 1)
    const int32_t* first_block;   // device array: first flat block of each pair (n > 1)
    int32_t n;
    int32_t total_blocks;
    int32_t cost_volume;          // batches (n > 1): every pair has a cost volume
};

// Raw full-resolution frames -> rectified half-resolution pair (calib::remap at the maps'
// resolution, then downsample2). side 0 is the left camera, side 1 the right.
struct RectifyDesc {
    const uint8_t* raw[2];
    const float* map[2];   // [map_h][map_w][2] (x, y) source coordinates, 16-byte aligned
    uint8_t* out[2];
    int64_t raw_pitch;     // bytes
    int64_t out_pitch;     // bytes
    int32_t raw_w, raw_h;
    int32_t map_w;         // map row length in entries (map_h >= 2 * out_h)
    int32_t out_w, out_h;  // map_w / 2, map_h / 2
};

// refine_level at one pyramid level (refine_kernels.hip).
struct RefineDesc {
    const uint8_t* left;
    const uint8_t* right;
    int64_t pitch;        // bytes, both images
    int32_t w, h;
    const float* coarse;  // median3 of the next coarser level's disparity, cw x ch
    int32_t cw, ch;
    float* disp;          // w x h outputs
    float* conf;
    int32_t radius, search_radius, reject_edge, subpixel;
    float min_zncc;
};

// refine_detail of a w x h disparity against the W x H full-resolution pair.
struct DetailDesc {
    const uint8_t* fl;
    const uint8_t* fr;
    int64_t fpitch;
    int32_t W, H;
    const float* first;   // input disparity
    float* out;           // refined copy
    int32_t w, h;
    int32_t radius, K, subpixel;
    float step, min_zncc;
};

// disparity_to_points over one w x h disparity (frame_kernels.hip): points and normals as float4
// (x, y, z, 1 where valid; zero elsewhere), weights.
struct PointsDesc {
    const float* disp;
    const float* conf;     // may be null: weight 1
    float* points;         // float4 [y][x]
    float* normals;        // float4 [y][x]
    float* weights;
    int32_t w, h;
    double f, cx, cy, baseline, cx_offset;
    float min_depth, max_depth, max_depth_jump, edge_weight, weight_reference_depth;
    float rc_x;            // the right camera centre's x as a float (the baseline)
};

// Host-side launchers (sgm_kernels.hip). All enqueue on `stream` and return the launch status.
hipError_t launch_census(const LaunchTable& t, const SgmConsts& c, hipStream_t stream);
hipError_t launch_cost_volume(const LaunchTable& t, const SgmConsts& c, hipStream_t stream);
// group: lanes per scanline, 16 (any dpl, 4/8/16 lines per block), 32 (dpl % 4 == 0, 4/8 lines) or 64
// (dpl % 16 == 0, 4 lines).
bool paths_group_ok(int dpl, int group, int lines_per_block);
hipError_t launch_paths(int dpl, bool padded, const LaunchTable& t, const SgmConsts& c, int lines_per_block, int group,
                        hipStream_t stream);
// threads: per row block, kBlock or 1024: the large block (1024 threads, 512 for the layouts where
// those would spill) for rows staged in LDS with byte path costs; other forms run kBlock whatever is
// asked.
hipError_t launch_rows(int dpl, bool padded, const LaunchTable& t, const SgmConsts& c, int max_w, hipStream_t stream,
                       int threads = kBlock);
hipError_t launch_median(const LaunchTable& t, hipStream_t stream);
hipError_t launch_cost8(int dpl, const LaunchTable& t, const SgmConsts& c, hipStream_t stream);
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream);
// Band sweeps of 8-path byte-cost batches (sweep_kernels.hip): band `band` of rows [band * band_rows,
// ...) of every pair of `t` (tile-column blocks per pair, sweep_blocks), both sweeps in one launch; and
// the row kernels that read the sweep records (rows_sweep.hip).
size_t sweep_lds_bytes(int dpl, int tile, int band_rows);
// A sweep block shape the kernel takes: tile in {16, 32, 48, 64}, a staged row within the block's
// loads, and at most lds_limit bytes of LDS.
bool sweep_shape_ok(int dpl, int tile, int band_rows, size_t lds_limit);
int sweep_blocks(int w, int tile);
hipError_t launch_sweep(int dpl, bool padded, const LaunchTable& t, const SgmConsts& c, int tile, int band_rows, int band,
                        hipStream_t stream);
hipError_t launch_rows_sweep(int dpl, bool padded, const LaunchTable& t, const SgmConsts& c, int max_w, hipStream_t stream);
hipError_t launch_rectify(const RectifyDesc& r, hipStream_t stream);

// Host-side launchers of the stages around SGM (refine_kernels.hip).
hipError_t launch_downsample2(const uint8_t* s0, const uint8_t* s1, int64_t spitch, int sw, int sh, uint8_t* d0, uint8_t* d1,
                              int64_t dpitch, hipStream_t stream);
hipError_t launch_median3(const float* in, float* out, int w, int h, hipStream_t stream);
hipError_t launch_refine(const RefineDesc& r, hipStream_t stream);
hipError_t launch_detail(const DetailDesc& d, hipStream_t stream);
hipError_t launch_occlusion(float* disp, int w, int h, float tol, hipStream_t stream);
hipError_t launch_spikes(const float* in, float* out, int w, int h, float thr, hipStream_t stream);
hipError_t launch_ccl(const float* disp, int32_t* labels, uint32_t* local, uint32_t* sizes, int w, int h, float max_diff,
                      hipStream_t stream);
hipError_t launch_finalize(const float* disp, const int32_t* labels, const uint32_t* sizes, int max_region,
                           const float* conf_src, float* out_disp, float* out_conf, int w, int h, hipStream_t stream);
hipError_t launch_rectify_full(const RectifyDesc& r, uint8_t* f0, uint8_t* f1, int64_t fpitch, int fw, int fh,
                               hipStream_t stream);

// Host-side launchers of the point image and the depth packing (frame_kernels.hip).
hipError_t launch_points(const PointsDesc& p, hipStream_t stream);
hipError_t launch_pack(const float* points, const float* weights, uint8_t* out, int w, int h, hipStream_t stream);

// Per-pair block counts for each stage (the launchers' grid geometry).
int census_blocks(int w, int h);
int path_blocks(const PairDesc& d, int lines_per_block);
int row_blocks(const PairDesc& d);
int cost8_blocks(int w, int h);
int median_blocks(int w, int h);

// Disparities per lane for a disparity count, or 0 when unsupported (> 512): the layout with the
// largest footprint, which allocations are sized by.
int dpl_for(int nd);
// The layout a w x h pair runs with (never wider than dpl_for).
int dpl_for_pair(int nd, int w, int h);
// Path costs fit a byte when cells + the largest step penalty (P2, or max(P2, P1 + 1) with adaptive P2)
// <= 255 and P1 <= 32767 (the byte kernels' 16-bit lane arithmetic adds P1 to a 0x7FFF sentinel);
// otherwise they are u16, added with saturation where a sum only competes in a minimum.
inline bool wide_path_costs(int cells, int p1, int p2, bool adaptive) {
    const int64_t pen = adaptive ? (p2 > (int64_t)p1 + 1 ? p2 : (int64_t)p1 + 1) : p2;
    return cells + pen > 255 || p1 > 32767;
}
// Bytes of step penalties per pixel: four axes, one byte or one u16 each.
inline size_t edge_bytes_per_px(bool wide) { return wide ? 8 : 4; }
// Census row layout for width w: x - min_d - k must address a word of the row for every pixel x
// and lane disparity k < dp, so the path kernel loads disparity windows without clamping.
inline void census_layout(int w, int min_d, int dp, int32_t& padl, int32_t& pitch) {
    padl = min_d + dp - 1 > 0 ? min_d + dp - 1 : 0;
    const int padr = min_d < 0 ? -min_d : 0;
    pitch = (padl + w + padr + 1) & ~1;
}
// Fills line_off for the 4 or 8 directions of the reference's schedule.
void fill_lines(PairDesc& d);

}  // namespace sgm
