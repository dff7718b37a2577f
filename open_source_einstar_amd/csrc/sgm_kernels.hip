// gfx950 (CDNA4) kernels of the SGM stereo hot path.
//
// Semantics follow the reference CPU matcher (libs/depth/src/stereo.cpp:44-96, 129-147, 328-346,
// 361-463); the layout and the parallel decomposition are this library's own:
//
//   census   one thread per pixel, LDS tile with halo                  -> census words u64
//   paths    one 16-lane DPP row (or two joined rows) per scanline,    -> L u8 [y][x][dir][dp]
//            all directions of all pairs in one launch; a lane holds
//            consecutive disparities as packed u16 pairs; the costs are
//            read from a u8 cost volume or recomputed from the census
//            words on the fly
//   rows     one block per image row: S = sum of the L_dir, left WTA   -> disparity, status, costs
//            with uniqueness + subpixel, right WTA by LDS atomic-min
//            of packed (S << 16 | k) keys, then the left/right check
//   median   3x3 median of valid neighbours                           -> disparity_median
//
// Every path value satisfies 0 <= L <= C + p2 <= 255 (the context enforces cells + p2 <= 255), so
// L fits a byte and S = sum over <= 8 paths <= 2040 never wraps: the direction order and the
// per-direction split of the sum do not change any result.
//
// This file holds the census, debug cost-volume, median, rectification and copy kernels and the
// launch dispatch; the path, cost-volume and row kernels are instantiated in paths_*.hip and
// rows_kernels.hip from sgm_kernels.h.
#include "sgm_kernels.h"

namespace sgm {
namespace {

// ------------------------------------------------------------------------------------ census
// census_transform, stereo.cpp:328-346: rows outer, columns inner, centre skipped, each compare
// shifted in at bit 0 (the first compare ends at bit cells-2); border pixels are 0.
//
// The left image's tiles also write each pixel's step penalties (stereo.cpp:62-65) for the four
// path axes r = (1, 0), (0, 1), (1, 1), (1, -1) as four bytes: byte k is the penalty of the step
// into this pixel from p - r_k, max(P1 + 1, P2 / (|dI|/8 + 1)) when adaptive, P2 otherwise (0
// where p - r_k leaves the image; no path steps from there). A step along -r_k uses byte k of the
// pixel it leaves. With byte path costs every penalty fits a byte (cells + penalty <= 255); with u16
// path costs (SgmConsts::wide_costs) the four penalties are u16 halves of two words.

constexpr int kTile = 16;  // median tiles
constexpr int kMaxRx = 4, kMaxRy = 3;
// A census block covers 64 x 16 pixels; each thread takes four adjacent pixels of one row (16
// threads a row, a wave four rows). The LDS tile holds the block's pixels with a halo of kMaxRx
// columns and kMaxRy rows, one u16 per pixel (0 outside the image), so that a 32-bit LDS word holds
// two horizontally neighbouring pixels.
constexpr int kCenW = 64, kCenH = 16, kCenPx = 4;
constexpr int kTileW = kCenW + 2 * kMaxRx, kTileH = kCenH + 2 * kMaxRy;
static_assert(kCenW / kCenPx * kCenH == kBlock, "one census thread per four pixels");

// 32 bits of the census words of pixels 2q and 2q + 1: byte k of `even` is byte 1 of t_k, byte k of
// `odd` is byte 3 of t_k (t_k holds bits 8k .. 8k + 7 of both words, relative to the 32-bit half).
__device__ __forceinline__ void census_gather(uint32_t t0, uint32_t t1, uint32_t t2, uint32_t t3, uint32_t& even,
                                              uint32_t& odd) {
    const uint32_t p01 = __builtin_amdgcn_perm(t1, t0, 0x07030501u);  // t0.b1 t1.b1 t0.b3 t1.b3
    const uint32_t p23 = __builtin_amdgcn_perm(t3, t2, 0x07030501u);
    even = __builtin_amdgcn_perm(p23, p01, 0x05040100u);
    odd = __builtin_amdgcn_perm(p23, p01, 0x07060302u);
}

// One 64x16 tile of one image (lb counts the left image's tiles, then the right's). Every thread of
// the block calls it (it synchronises once).
template <int RX, int RY>  // the census radii when known at compile time, else -1
__device__ __forceinline__ void census_body(const PairDesc& P, const SgmConsts& c, int lb, SGM_LDS uint16_t* tile,
                                            SGM_LDS uint32_t* p2q) {
    const int tx_n = (P.w + kCenW - 1) / kCenW, ty_n = (P.h + kCenH - 1) / kCenH;
    const int side = lb / (tx_n * ty_n);
    const int tl = lb - side * tx_n * ty_n;
    const int x0 = (tl % tx_n) * kCenW, y0 = (tl / tx_n) * kCenH;
    const uint8_t* img = side ? P.right : P.left;
    uint64_t* out = side ? P.census_r : P.census_l;
    // Every halo byte is fetched before the first one is written to LDS.
    constexpr int kTileN = kTileW * kTileH;
    constexpr int kLoads = (kTileN + kBlock - 1) / kBlock;
    uint32_t px[kLoads];
#pragma unroll
    for (int j = 0; j < kLoads; ++j) {
        const int i = (int)threadIdx.x + j * kBlock;
        const int ly = i / kTileW, lx = i - ly * kTileW;
        const int gx = x0 - kMaxRx + lx, gy = y0 - kMaxRy + ly;
        px[j] = (i < kTileN && gx >= 0 && gy >= 0 && gx < P.w && gy < P.h) ? img[(int64_t)gy * P.pitch + gx] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kLoads; ++j) {
        const int i = (int)threadIdx.x + j * kBlock;
        if (i < kTileN) tile[i] = (uint16_t)px[j];
    }
    if (threadIdx.x < 32) p2q[threadIdx.x] = (uint32_t)(c.adaptive_p2 ? c.p2_by_q[threadIdx.x] : c.p2);
    __syncthreads();
    const int lx = (int)(threadIdx.x % (kCenW / kCenPx)) * kCenPx, ly = (int)(threadIdx.x / (kCenW / kCenPx));
    const int x = x0 + lx, y = y0 + ly;
    if (x >= P.w || y >= P.h) return;
    const int rx = c.census_rx, ry = c.census_ry;
    const SGM_LDS uint16_t* ctr = &tile[(ly + kMaxRy) * kTileW + lx + kMaxRx];  // pixel (x, y)
    uint64_t bits[kCenPx];
    if constexpr (RX >= 0) {
        // Compile-time window, two pixels per 32-bit word: the packed u16 difference neighbour -
        // centre has its half's high byte 0xFF exactly when the neighbour is below the centre, and
        // compare i of a pixel lands at bit cells - 2 - i of its word. T[q][k] collects bits 8k ..
        // 8k + 7 of pixels 2q (byte 1) and 2q + 1 (byte 3).
        constexpr int kCells = (2 * RX + 1) * (2 * RY + 1);
        constexpr int kOct = (kCells - 1 + 7) / 8;
        uint32_t T[2][8] = {};
        // Word j of a tile row: pixels x - 4 + 2j and x - 3 + 2j (tile column lx + 2j is even).
        const SGM_LDS uint32_t* row0 = reinterpret_cast<const SGM_LDS uint32_t*>(ctr - kMaxRx);
        const uint32_t C0 = row0[2], C1 = row0[3];  // centres: pixels (x, x + 1), (x + 2, x + 3)
#pragma unroll
        for (int v = -RY; v <= RY; ++v) {
            const SGM_LDS uint32_t* rw = row0 + v * (kTileW / 2);
            uint32_t wd[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) wd[j] = rw[j];
#pragma unroll
            for (int u = -RX; u <= RX; ++u) {
                if (u == 0 && v == 0) continue;
                const int i = (v + RY) * (2 * RX + 1) + (u + RX) - ((v > 0 || (v == 0 && u > 0)) ? 1 : 0);
                const int pos = kCells - 2 - i;
                const uint32_t m = 0x01000100u << (pos & 7);
                uint32_t a0, a1;  // neighbours (u, v) of pixels (x, x + 1) and (x + 2, x + 3)
                if ((u & 1) == 0) {
                    a0 = wd[(u + 4) / 2];
                    a1 = wd[(u + 4) / 2 + 1];
                } else {
                    a0 = __builtin_amdgcn_alignbit(wd[(u + 5) / 2], wd[(u + 3) / 2], 16);
                    a1 = __builtin_amdgcn_alignbit(wd[(u + 5) / 2 + 1], wd[(u + 3) / 2 + 1], 16);
                }
                T[0][pos >> 3] |= pk_sub(a0, C0) & m;
                T[1][pos >> 3] |= pk_sub(a1, C1) & m;
            }
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            uint32_t lo_e, lo_o, hi_e = 0, hi_o = 0;
            census_gather(T[q][0], T[q][1], T[q][2], T[q][3], lo_e, lo_o);
            if constexpr (kOct > 4) census_gather(T[q][4], T[q][5], T[q][6], T[q][7], hi_e, hi_o);
            bits[2 * q] = ((uint64_t)hi_e << 32) | lo_e;
            bits[2 * q + 1] = ((uint64_t)hi_o << 32) | lo_o;
        }
    } else {
        // rows outer, columns inner, each compare shifted in at bit 0
#pragma unroll
        for (int p = 0; p < kCenPx; ++p) {
            const uint32_t center = ctr[p];
            uint64_t b = 0;
            for (int v = -ry; v <= ry; ++v) {
                const SGM_LDS uint16_t* row = ctr + p + v * kTileW - rx;
                for (int u = 0; u <= 2 * rx; ++u) {
                    if (u == rx && v == 0) continue;
                    b = (b << 1) | (uint64_t)(row[u] < center);
                }
            }
            bits[p] = b;
        }
    }
    const int64_t orow = (int64_t)y * P.census_pitch + P.census_padl;
    const bool row_in = y >= ry && y < P.h - ry;
#pragma unroll
    for (int p = 0; p < kCenPx; ++p) {
        const int xp = x + p;
        if (xp < P.w) out[orow + xp] = (row_in && xp >= rx && xp < P.w - rx) ? bits[p] : 0ull;
    }
    if (side == 0) {
#pragma unroll
        for (int p = 0; p < kCenPx; ++p) {
            const int xp = x + p;
            if (xp >= P.w) break;
            const SGM_LDS uint16_t* q = ctr + p;
            const int I = q[0];
            auto pen = [&](int v, bool inside) { return inside ? (uint32_t)p2q[abs(I - v) >> 3] : 0u; };
            const uint32_t e0 = pen(q[-1], xp > 0), e1 = pen(q[-kTileW], y > 0), e2 = pen(q[-kTileW - 1], xp > 0 && y > 0),
                           e3 = pen(q[kTileW - 1], xp > 0 && y + 1 < P.h);
            if (c.wide_costs) {
                // u16 penalties (saturated: a step penalty only competes in a minimum against values
                // below 65536), axes 0-1 then 2-3.
                auto s16 = [](uint32_t v) { return v > 0xFFFFu ? 0xFFFFu : v; };
                reinterpret_cast<u32x2*>(P.edge)[(int64_t)y * P.w + xp] =
                    u32x2{s16(e0) | (s16(e1) << 16), s16(e2) | (s16(e3) << 16)};
            } else {
                P.edge[(int64_t)y * P.w + xp] = e0 | (e1 << 8) | (e2 << 16) | (e3 << 24);
            }
        }
    }
}

template <int RX, int RY>
__global__ __launch_bounds__(kBlock) void census_kernel(LaunchTable t, SgmConsts c) {
    __shared__ __attribute__((aligned(16))) uint16_t tile[kTileH * kTileW];
    __shared__ uint32_t p2q[32];
    int pi, lb;
    locate(t, blockIdx.x, pi, lb);
    const PairDesc P = load_desc(t, pi);
    census_body<RX, RY>(P, c, lb, lds(tile), lds(p2q));
}

// ------------------------------------------------------------------------------- cost volume
// Debug output only (the path kernel recomputes costs): the cost fill of stereo.cpp:366-380.

__global__ __launch_bounds__(kBlock) void cost_volume_kernel(LaunchTable t, SgmConsts c) {
    int pi, lb;
    locate(t, blockIdx.x, pi, lb);
    const PairDesc P = load_desc(t, pi);
    const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
    if (i >= (int64_t)P.w * P.h) return;
    const int y = (int)(i / P.w), x = (int)(i - (int64_t)y * P.w);
    const int64_t row = (int64_t)y * P.census_pitch + P.census_padl;
    const uint64_t a = P.census_l[row + x];
    uint16_t* dst = P.cost_volume + i * P.nd;
    for (int k = 0; k < P.nd; ++k) {
        const int xr = x - (P.min_d + k);
        dst[k] = (xr >= 0 && xr < P.w) ? (uint16_t)__popcll(a ^ P.census_r[row + xr]) : (uint16_t)c.invalid_cost;
    }
}

// ------------------------------------------------------------------------------------ median
// median3, stereo.cpp:129-147: the n/2-th smallest (0-based) of the valid in-image 3x3 values.

// Sorts 9 values ascending with a 25-comparator network (depth 9; it sorts every 0/1 input, so every
// input). Any sorting network gives the same values in the same order.
__device__ __forceinline__ void sort9(float (&b)[9]) {
    constexpr int kNet[25][2] = {{0, 3}, {1, 7}, {2, 5}, {4, 8}, {0, 7}, {2, 4}, {3, 8}, {5, 6}, {0, 2},
                                 {1, 3}, {4, 5}, {7, 8}, {1, 4}, {3, 6}, {5, 7}, {0, 1}, {2, 4}, {3, 5},
                                 {6, 8}, {2, 3}, {4, 5}, {6, 7}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
    for (int c = 0; c < 25; ++c) {
        const float lo = fminf(b[kNet[c][0]], b[kNet[c][1]]), hi = fmaxf(b[kNet[c][0]], b[kNet[c][1]]);
        b[kNet[c][0]] = lo;
        b[kNet[c][1]] = hi;
    }
}

// One 16x16 tile of the disparity map; its pixels and a one-pixel halo are read once into LDS
// (invalid outside the image). Every thread of the block calls it (it synchronises once).
__device__ __forceinline__ void median_body(const PairDesc& P, int lb, SGM_LDS float* tile) {
    constexpr int kT = kTile + 2;
    const int tx_n = (P.w + kTile - 1) / kTile;
    const int x0 = (lb % tx_n) * kTile, y0 = (lb / tx_n) * kTile;
    const float* in = P.disp;
    for (int i = threadIdx.x; i < kT * kT; i += kBlock) {
        const int ly = i / kT, lx = i - ly * kT, gx = x0 - 1 + lx, gy = y0 - 1 + ly;
        tile[i] = (gx >= 0 && gy >= 0 && gx < P.w && gy < P.h) ? in[(int64_t)gy * P.w + gx] : kInvalid;
    }
    __syncthreads();
    const int lx = threadIdx.x & (kTile - 1), ly = threadIdx.x >> 4;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= P.w || y >= P.h) return;
    const SGM_LDS float* c = tile + (ly + 1) * kT + lx + 1;
    float out = kInvalid;
    if (*c > -1.0e8f) {
        float buf[9];
        int n = 0;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                const float q = c[dy * kT + dx];
                // Invalid or outside: +inf sorts after every valid value and is never selected.
                const bool ok = q > -1.0e8f;
                buf[dy * 3 + dx + 4] = ok ? q : __builtin_inff();
                n += ok ? 1 : 0;
            }
        sort9(buf);  // valid values ascending, +inf last
        float sel = buf[0];
#pragma unroll
        for (int i = 1; i < 9; ++i) sel = (n / 2) == i ? buf[i] : sel;
        out = sel;
    }
    P.median[(int64_t)y * P.w + x] = out;
}

__global__ __launch_bounds__(kBlock) void median_kernel(LaunchTable t) {
    __shared__ float tile[(kTile + 2) * (kTile + 2)];
    int pi, lb;
    locate(t, blockIdx.x, pi, lb);
    const PairDesc P = load_desc(t, pi);
    median_body(P, lb, lds(tile));
}

// ------------------------------------------------------------------------------ rectification
// calib::remap (libs/calib/src/rectify.cpp:133-148) at the maps' resolution followed by
// downsample2 (stereo.cpp:348-359), fused: a thread makes one output pixel from its 2x2 block of
// full-resolution samples, reading their four (x, y) map entries as two 16-byte loads, and the
// full-resolution image never exists. The bilinear expression keeps the reference's operation
// order and the library is built with -ffp-contract=off, so every sample rounds as on the CPU.

typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kRectW = 64, kRectH = kBlock / kRectW;

__device__ __forceinline__ uint32_t remap_sample(const SGM_GLOBAL uint8_t* src, int64_t pitch, int sw, int sh, float sx,
                                                 float sy) {
    // floor(s) in [0, size - 2]  <=>  0 <= s < size - 1 for every finite s; NaN fails both tests.
    // Branch-free: a sample outside reads the 2x2 pixels at (0, 0) and its value is replaced by 0. Each
    // row's two taps come in one 16-bit load (unaligned where the left pixel's column is odd). A frame
    // narrower or lower than 2 pixels has no sample inside, and no 2x2 block to read (a uniform test).
    if (sw < 2 || sh < 2) return 0;
    const bool in = sx >= 0.0f && sy >= 0.0f && sx < (float)(sw - 1) && sy < (float)(sh - 1);
    const float flx = __builtin_floorf(sx), fly = __builtin_floorf(sy);
    const float fx = sx - flx, fy = sy - fly;
    const int ix = in ? (int)flx : 0, iy = in ? (int)fly : 0;
    const SGM_GLOBAL uint8_t* p = src + (int64_t)iy * pitch + ix;
    typedef uint16_t u16u __attribute__((aligned(1)));
    const uint32_t top = *reinterpret_cast<const SGM_GLOBAL u16u*>(p);
    const uint32_t bot = *reinterpret_cast<const SGM_GLOBAL u16u*>(p + pitch);
    const float a = (float)(top & 0xFFu), b = (float)(top >> 8), c = (float)(bot & 0xFFu), d = (float)(bot >> 8);
    const float v = (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy;
    return in ? (uint32_t)__builtin_fminf(v + 0.5f, 255.0f) : 0u;  // v >= 0: the lower clamp never binds
}

__global__ __launch_bounds__(kBlock) void rectify_kernel(RectifyDesc r) {
    const int side = (int)blockIdx.z;
    const int x = (int)blockIdx.x * kRectW + (int)threadIdx.x;
    const int y = (int)blockIdx.y * kRectH + (int)threadIdx.y;
    if (x >= r.out_w || y >= r.out_h) return;
    const SGM_GLOBAL uint8_t* src = (const SGM_GLOBAL uint8_t*)r.raw[side];
    const SGM_GLOBAL float* map = (const SGM_GLOBAL float*)r.map[side] + ((int64_t)(2 * y) * r.map_w + 2 * x) * 2;
    const f32x4 m0 = *(const SGM_GLOBAL f32x4*)map;
    const f32x4 m1 = *(const SGM_GLOBAL f32x4*)(map + 2 * (int64_t)r.map_w);
    const uint32_t acc = remap_sample(src, r.raw_pitch, r.raw_w, r.raw_h, m0.x, m0.y) +
                         remap_sample(src, r.raw_pitch, r.raw_w, r.raw_h, m0.z, m0.w) +
                         remap_sample(src, r.raw_pitch, r.raw_w, r.raw_h, m1.x, m1.y) +
                         remap_sample(src, r.raw_pitch, r.raw_w, r.raw_h, m1.z, m1.w);
    ((SGM_GLOBAL uint8_t*)r.out[side])[(int64_t)y * r.out_pitch + x] = (uint8_t)((acc + 2) >> 2);
}

// -------------------------------------------------------------------------------------- copy
// Plain 16-byte-per-lane copy: the known-size reference for HBM traffic counters.

__global__ __launch_bounds__(kBlock) void copy_kernel(u32x4* __restrict__ dst, const u32x4* __restrict__ src, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += (int64_t)gridDim.x * kBlock) dst[i] = src[i];
}

// -------------------------------------------------------------------------- launch dispatch

}  // namespace

// 6 disparities per lane is never chosen: its lanes read and write path costs in 2-byte pieces and ran
// slower than 8 per lane with the pad masked at every size measured (8 paths, D = 96: path kernel
// 1.00 against 0.49 ms at 1280x1024, 120 against 56 us at 320x256). 2 per lane has the same pieces;
// it is kept for pairs of at most kSmallPairPx pixels, where the walks are short and their latency
// rather than the bytes sets the time (D = 32: 34 against 39 us at 320x256, 129 against 103 us at
// 640x512 for 2 and 4 per lane). dpl_for gives the layout with the largest footprint for a disparity
// count (what allocations are sized by), dpl_for_pair the one a pair runs with.
constexpr int64_t kSmallPairPx = 256 * 1024;

int dpl_for(int nd) {
    if (nd <= 0) return 0;
    const int need = (nd + kGroup - 1) / kGroup;
    for (const int d : {4, 8, 12, 16, 24, 32})
        if (d >= need) return d;
    return 0;
}

int dpl_for_pair(int nd, int w, int h) {
    if (nd > 0 && nd <= 2 * kGroup && (int64_t)w * h <= kSmallPairPx) return 2;
    return dpl_for(nd);
}

void fill_lines(PairDesc& d) {
    int total = 0;
    for (int i = 0; i < kMaxDirs; ++i) {
        d.line_off[i] = total;
        if (i < d.ndirs) total += path_lines(i, d.w, d.h);
    }
    d.total_lines = total;
}

int census_blocks(int w, int h) { return 2 * ((w + kCenW - 1) / kCenW) * ((h + kCenH - 1) / kCenH); }
int path_blocks(const PairDesc& d, int lines_per_block) { return (d.total_lines + lines_per_block - 1) / lines_per_block; }
int row_blocks(const PairDesc& d) { return d.h; }
int cost8_blocks(int w, int h) { return h * cost8_segments(w); }
int median_blocks(int w, int h) { return ((w + kTile - 1) / kTile) * ((h + kTile - 1) / kTile); }

hipError_t launch_census(const LaunchTable& t, const SgmConsts& c, hipStream_t stream) {
    if (t.total_blocks <= 0) return hipSuccess;
    // The windows of the reference's parameter sets (9x7 and 5x5) are compiled in; others run the
    // generic loop.
    if (c.census_rx == 4 && c.census_ry == 3)
        hipLaunchKernelGGL((census_kernel<4, 3>), dim3(t.total_blocks), dim3(kBlock), 0, stream, t, c);
    else if (c.census_rx == 2 && c.census_ry == 2)
        hipLaunchKernelGGL((census_kernel<2, 2>), dim3(t.total_blocks), dim3(kBlock), 0, stream, t, c);
    else
        hipLaunchKernelGGL((census_kernel<-1, -1>), dim3(t.total_blocks), dim3(kBlock), 0, stream, t, c);
    return hipGetLastError();
}

hipError_t launch_cost_volume(const LaunchTable& t, const SgmConsts& c, hipStream_t stream) {
    if (t.total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(cost_volume_kernel, dim3(t.total_blocks), dim3(kBlock), 0, stream, t, c);
    return hipGetLastError();
}

bool paths_group_ok(int dpl, int group, int lines_per_block) {
    if (group == 16) return lines_per_block == 4 || lines_per_block == 8 || lines_per_block == 16;
    if (group == 64) return dpl % 16 == 0 && lines_per_block == 4;
    return group == 32 && dpl % 4 == 0 && (lines_per_block == 4 || lines_per_block == 8);
}

hipError_t launch_paths(int dpl, bool padded, const LaunchTable& t, const SgmConsts& c, int lines_per_block, int group,
                        hipStream_t stream) {
    if (t.total_blocks <= 0) return hipSuccess;
    if (c.wide_costs) group = 16;  // u16 path costs: 16 lanes per scanline only
    if (!paths_group_ok(dpl, group, lines_per_block)) return hipErrorInvalidValue;
    const PathLaunch pl{padded, (t.n == 1 ? t.one.cvol : nullptr) != nullptr || t.cost_volume != 0, c.slant_steps != 0,
                        (unsigned)(lines_per_block * group)};
    if (c.wide_costs) return launch_paths_w16(dpl, pl, t, c, stream);
    if (group == 16) return dpl <= 8 ? launch_paths_g16_lo(dpl, pl, t, c, stream) : launch_paths_g16_hi(dpl, pl, t, c, stream);
    if (group == 32) return launch_paths_g32(dpl, pl, t, c, stream);
    return launch_paths_g64(dpl, pl, t, c, stream);
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream) {
    const int64_t n16 = (int64_t)(bytes / 16);
    if (n16 <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n16 + kBlock - 1) / kBlock, 256 * 64);
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, stream, static_cast<u32x4*>(dst),
                       static_cast<const u32x4*>(src), n16);
    return hipGetLastError();
}

hipError_t launch_rectify(const RectifyDesc& r, hipStream_t stream) {
    if (r.out_w <= 0 || r.out_h <= 0) return hipSuccess;
    const dim3 grid((unsigned)((r.out_w + kRectW - 1) / kRectW), (unsigned)((r.out_h + kRectH - 1) / kRectH), 2);
    hipLaunchKernelGGL(rectify_kernel, grid, dim3(kRectW, kRectH), 0, stream, r);
    return hipGetLastError();
}

hipError_t launch_median(const LaunchTable& t, hipStream_t stream) {
    if (t.total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(median_kernel, dim3(t.total_blocks), dim3(kBlock), 0, stream, t);
    return hipGetLastError();
}

}  // namespace sgm
