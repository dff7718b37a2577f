// Band sweeps: the six non-horizontal path directions of an 8-path batch, aggregated row by row so
// that each sweep writes one u16 partial sum per pixel and disparity instead of three u8 path-cost
// planes (sgm_kernels.h has the line kernels these replace for batches).
//
// The down sweep walks rows top to bottom with directions 2 (0,+1), 4 (+1,+1) and 7 (-1,+1); the up
// sweep walks bottom to top with 3 (0,-1), 6 (+1,-1) and 5 (-1,-1) (stereo.cpp:387-388 numbering).
// In both, each direction's predecessor of pixel (x, y) lies on the previous swept row at x, x - 1 or
// x + 1 ("vertical", "left" and "right" kinds below), and a pixel whose predecessor leaves the image
// starts its line (stereo.cpp:391-402: every pixel lies on exactly one line per direction, so this
// is the recurrence of the line walk, stereo.cpp:44-89, evaluated in another order).
//
// One launch covers one band of rows of every pair of a lane class; a block owns a tile of T <= 64
// columns of a pair. It also computes what its rows need from outside the tile -- the left kind up
// to (rows left in the band) columns to its left, the right kind as many to its right -- so blocks
// never wait for one another; the band's first row continues from the previous band's last row,
// which every block leaves in a seed buffer for its tile. State passes from row to row through LDS
// as bytes (with byte path costs 0 <= L <= cells + P2 <= 255), double buffered: one barrier a row.
//
// Per pixel the three directions' path costs are summed in registers (at most 3 * 255) and stored
// as u16 into the pixel's record [S_down[dp] u16, S_up[dp] u16, L0[dp] u8, L1[dp] u8]; the two
// horizontal directions come from the line kernel, and the row kernel adds the four parts.
#include "sgm_kernels.h"

namespace sgm {
namespace {

constexpr int kSweepCols = 4;  // core columns per 16-lane group: tiles of at most 64 columns
constexpr int kKinds = 3;

// Direction of each kind: 0 vertical, 1 predecessor at x - 1, 2 predecessor at x + 1.
__device__ __forceinline__ int sweep_dir(bool up, int kind) {
    return up ? (kind == 0 ? 3 : (kind == 1 ? 6 : 5)) : (kind == 0 ? 2 : (kind == 1 ? 4 : 7));
}

// A lane's DPL path-cost bytes (2-byte aligned at least) as (DPL + 3) / 4 words.
template <int DPL>
__device__ __forceinline__ void lds_load_bytes(const SGM_LDS uint8_t* p, uint32_t (&w)[(DPL + 3) / 4]) {
    if constexpr (DPL % 8 == 0) {
#pragma unroll
        for (int i = 0; i < DPL / 8; ++i) {
            const u32x2 q = reinterpret_cast<const SGM_LDS u32x2*>(p)[i];
            w[2 * i] = q.x;
            w[2 * i + 1] = q.y;
        }
    } else if constexpr (DPL % 4 == 0) {
#pragma unroll
        for (int i = 0; i < DPL / 4; ++i) w[i] = reinterpret_cast<const SGM_LDS uint32_t*>(p)[i];
    } else {
#pragma unroll
        for (int i = 0; i < DPL / 2; ++i) {
            const uint32_t h = reinterpret_cast<const SGM_LDS uint16_t*>(p)[i];
            w[i / 2] = (i & 1) ? (w[i / 2] | (h << 16)) : h;
        }
    }
}
template <int DPL>
__device__ __forceinline__ void lds_store_bytes(SGM_LDS uint8_t* p, const Packed<DPL>& b) {
    if constexpr (DPL % 8 == 0) {
#pragma unroll
        for (int i = 0; i < DPL / 8; ++i) reinterpret_cast<SGM_LDS u32x2*>(p)[i] = u32x2{b.w[2 * i], b.w[2 * i + 1]};
    } else if constexpr (DPL % 4 == 0) {
#pragma unroll
        for (int i = 0; i < DPL / 4; ++i) reinterpret_cast<SGM_LDS uint32_t*>(p)[i] = b.w[i];
    } else {
#pragma unroll
        for (int i = 0; i < DPL / 4; ++i) {
            reinterpret_cast<SGM_LDS uint16_t*>(p)[2 * i] = (uint16_t)b.w[i];
            reinterpret_cast<SGM_LDS uint16_t*>(p)[2 * i + 1] = (uint16_t)(b.w[i] >> 16);
        }
        reinterpret_cast<SGM_LDS uint16_t*>(p)[DPL / 2 - 1] = (uint16_t)b.w[DPL / 4];
    }
}
// Global memory, ordinary loads (the cost volume and the seeds stay cacheable).
template <int DPL>
__device__ __forceinline__ void global_load_bytes(const SGM_GLOBAL uint8_t* p, uint32_t (&w)[(DPL + 3) / 4]) {
    if constexpr (DPL % 16 == 0) {
#pragma unroll
        for (int i = 0; i < DPL / 16; ++i) {
            const u32x4 q = reinterpret_cast<const SGM_GLOBAL u32x4*>(p)[i];
            w[4 * i] = q.x;
            w[4 * i + 1] = q.y;
            w[4 * i + 2] = q.z;
            w[4 * i + 3] = q.w;
        }
    } else if constexpr (DPL % 8 == 0) {
#pragma unroll
        for (int i = 0; i < DPL / 8; ++i) {
            const u32x2 q = reinterpret_cast<const SGM_GLOBAL u32x2*>(p)[i];
            w[2 * i] = q.x;
            w[2 * i + 1] = q.y;
        }
    } else if constexpr (DPL % 4 == 0) {
#pragma unroll
        for (int i = 0; i < DPL / 4; ++i) w[i] = reinterpret_cast<const SGM_GLOBAL uint32_t*>(p)[i];
    } else {
#pragma unroll
        for (int i = 0; i < DPL / 2; ++i) {
            const uint32_t h = reinterpret_cast<const SGM_GLOBAL uint16_t*>(p)[i];
            w[i / 2] = (i & 1) ? (w[i / 2] | (h << 16)) : h;
        }
    }
}
// Bytes -> packed u16 pairs (low half = even disparity).
template <int DPL>
__device__ __forceinline__ void unpack_bytes(const uint32_t (&w)[(DPL + 3) / 4], uint32_t (&v)[DPL / 2]) {
#pragma unroll
    for (int i = 0; i < DPL / 4; ++i) {
        v[2 * i] = __builtin_amdgcn_perm(0u, w[i], 0x0C010C00u);
        v[2 * i + 1] = __builtin_amdgcn_perm(0u, w[i], 0x0C030C02u);
    }
    if constexpr (DPL % 4 == 2) {
        const uint32_t q = w[DPL / 4];
        v[DPL / 2 - 1] = (q & 0xFFu) | ((q & 0xFF00u) << 8);
    }
}

// The LDS of a sweep block, in this order:
//   state   two buffers (row parity) of sweep_slots(T, B) slots, each a pixel's dp path-cost bytes:
//           the vertical kind's T tile columns [x0, x0 + T), the left kind's T + B columns from x0 - B,
//           the right kind's T + B columns from x0;
//   minima  the slots' line minima (u32, both halves), two buffers;
//   costs   two buffers (row parity) of the row's matching costs, T + 2B columns from x0 - B;
//   pens    two buffers of the row's step-penalty words and of its predecessor row's, T + 2B + 1
//           columns from x0 - B each.
// The next row's costs and penalty words are loaded by the whole block while the current row is
// computed, and written to LDS at the row's end, so no item waits for global memory; the band's
// first row finds its predecessors (the previous band's last row, from the seeds) in state buffer 1.
__host__ __device__ inline int sweep_slots(int tile, int band) { return 3 * tile + 2 * band; }
__host__ __device__ inline int sweep_cols(int tile, int band) { return tile + 2 * band; }
__device__ __forceinline__ int sweep_slot(int kind, int x, int x0, int tile, int band) {
    return kind == 0 ? x - x0 : (kind == 1 ? tile + (x - x0 + band) : 2 * tile + band + (x - x0));
}
// 16-byte loads per thread that stage one row's costs (T + 2B columns of dp bytes).
__host__ __device__ constexpr int sweep_stage_loads(int dpl) { return dpl <= 8 ? 4 : (dpl <= 16 ? 6 : 8); }

template <int DPL, bool PAD, bool SLANT, bool UP>
__device__ __forceinline__ void sweep_body(const LaunchTable& t, const SgmConsts& c, int tile, int band_rows, int band,
                                           SGM_LDS uint32_t* smem_words) {
    constexpr int R = DPL / 2;
    constexpr int CW = (DPL + 3) / 4;
    constexpr int NL = sweep_stage_loads(DPL);
    int pi, ti;
    locate(t, blockIdx.x, pi, ti);
    const PairDesc P = load_desc(t, pi);
    const int W = P.w, H = P.h, dp = P.dp;
    const int t0 = band * band_rows;
    if (t0 >= H) return;  // the pair has fewer bands (block-uniform: no barrier reached)
    const int nrows = min(band_rows, H - t0);
    const int x0 = ti * tile, x1 = min(W, x0 + tile);
    const int xb = x0 - band_rows;  // column of staged index 0
    const int ns = sweep_slots(tile, band_rows), ncol = sweep_cols(tile, band_rows);
    SGM_LDS uint8_t* lbytes = (SGM_LDS uint8_t*)smem_words;
    SGM_LDS uint32_t* lmin = (SGM_LDS uint32_t*)(lbytes + (size_t)2 * ns * dp);
    SGM_LDS uint8_t* lcost = (SGM_LDS uint8_t*)(lmin + 2 * ns);
    SGM_LDS uint32_t* lpen = (SGM_LDS uint32_t*)(lcost + (size_t)2 * ncol * dp);  // [parity][row, pred row][ncol + 1]

    const int grp = threadIdx.x >> 4, g = threadIdx.x & (kGroup - 1);
    const int k0 = g * DPL;
    const uint32_t p1p1 = (uint32_t)c.p1 * 0x10001u;
    uint32_t padm[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        padm[r] = PAD ? ((k0 + 2 * r >= P.nd ? 0x7FFFu : 0u) | (k0 + 2 * r + 1 >= P.nd ? 0x7FFFu << 16 : 0u)) : 0u;
    const LineEdge edge{};  // unused with 16 lanes per line
    const int64_t wdp = (int64_t)W * dp;
    // Seeds of this sweep: [parity][kind][x][dp] bytes.
    SGM_GLOBAL uint8_t* seeds = gptr(P.seed) + (UP ? 2 : 0) * kKinds * wdp;
    const SGM_GLOBAL uint8_t* seed_in = seeds + ((band + 1) & 1) * kKinds * wdp;
    SGM_GLOBAL uint8_t* seed_out = seeds + (band & 1) * kKinds * wdp;
    const SGM_GLOBAL uint8_t* cvol = gptr(P.cvol);
    const SGM_GLOBAL uint32_t* pen = gptr(P.edge);
    SGM_GLOBAL uint8_t* rec = gptr(P.lvol) + (UP ? 2 * dp : 0) + 2 * k0;  // this sweep's u16 sums
    const int64_t pxs = P.lvol_px_stride;

    // Row j of the band: image row, predecessor row, and the columns it computes, [lo, hi).
    auto row_y = [&](int j) { return UP ? H - 1 - (t0 + j) : t0 + j; };
    auto row_lo = [&](int j) { return max(0, x0 - (nrows - 1 - j)); };
    auto row_hi = [&](int j) { return min(W, x1 + (nrows - 1 - j)); };
    // A row's staging loads: its costs over [lo, hi) in 16-byte pieces, its penalty words over
    // [lo, hi) and its predecessor row's over [lo - 1, hi + 1).
    struct Stage {
        u32x4 q[NL];
        uint32_t e0, e1;
    };
    auto stage_load = [&](int j, Stage& st) {
        const int y = row_y(j), yp = UP ? y + 1 : y - 1;
        const int lo = row_lo(j), hi = row_hi(j);
        const int n16 = (hi - lo) * dp / 16;
        const SGM_GLOBAL u32x4* src = reinterpret_cast<const SGM_GLOBAL u32x4*>(cvol + ((int64_t)y * W + lo) * dp);
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int i = q * kBlock + (int)threadIdx.x;
            if (i < n16) st.q[q] = src[i];
        }
        const int x = lo + (int)threadIdx.x;
        if (x < hi) st.e0 = pen[(int64_t)y * W + x];
        const int xq = max(0, lo - 1) + (int)threadIdx.x;
        if (yp >= 0 && yp < H && xq < min(W, hi + 1)) st.e1 = pen[(int64_t)yp * W + xq];
    };
    auto stage_store = [&](int j, const Stage& st) {
        const int y = row_y(j), yp = UP ? y + 1 : y - 1;
        const int lo = row_lo(j), hi = row_hi(j);
        const int n16 = (hi - lo) * dp / 16;
        SGM_LDS u32x4* dst = reinterpret_cast<SGM_LDS u32x4*>(lcost + ((size_t)(j & 1) * ncol + (lo - xb)) * dp);
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int i = q * kBlock + (int)threadIdx.x;
            if (i < n16) dst[i] = st.q[q];
        }
        SGM_LDS uint32_t* pe = lpen + (size_t)(j & 1) * 2 * (ncol + 1);
        const int x = lo + (int)threadIdx.x;
        if (x < hi) pe[x - xb] = st.e0;
        const int xq = max(0, lo - 1) + (int)threadIdx.x;
        if (yp >= 0 && yp < H && xq < min(W, hi + 1)) pe[ncol + 1 + (xq - xb)] = st.e1;
    };

    // One step of kind `kind` at column x of band row j, in two phases so that a group can issue the
    // LDS loads of several items before it computes any: gather loads the predecessor's path costs,
    // line minimum and step-penalty word (nothing where the line starts), finish computes the new
    // path costs into v and stores them.
    struct Item {
        uint32_t pw[CW];
        uint32_t pm, E;
        bool start;
    };
    auto gather = [&](int kind, int x, int j, Item& it) {
        const int xp = x + (kind == 1 ? -1 : (kind == 2 ? 1 : 0));
        it.start = t0 + j == 0 || xp < 0 || xp >= W;
        if (!it.start) {
            const int slot_r = ((j - 1) & 1) * ns + sweep_slot(kind, xp, x0, tile, band_rows);
            lds_load_bytes<DPL>(lbytes + (size_t)slot_r * dp + k0, it.pw);
            it.pm = lmin[slot_r];
            // Step penalty (sgm_kernels.h, paths_body): byte dir/2 of the word of the pixel stepped
            // into for directions along their axis, of the pixel stepped from otherwise.
            const int dir = sweep_dir(UP, kind);
            const SGM_LDS uint32_t* pe = lpen + (size_t)(j & 1) * 2 * (ncol + 1);
            it.E = (dir & 1) == 0 ? pe[x - xb] : pe[ncol + 1 + (xp - xb)];
        }
    };
    auto load_cost = [&](int x, int j, uint32_t (&cw)[CW]) {
        lds_load_bytes<DPL>(lcost + ((size_t)(j & 1) * ncol + (x - xb)) * dp + k0, cw);
    };
    auto finish = [&](int kind, int x, int j, const uint32_t (&cw)[CW], const Item& it, uint32_t (&v)[R]) {
        uint32_t cc[R];
        unpack_bytes<DPL>(cw, cc);
        uint32_t mn = 0xFFFFFFFFu;
        if (it.start) {  // the line starts here
#pragma unroll
            for (int r = 0; r < R; ++r) {
                v[r] = PAD ? pk_max(cc[r], padm[r]) : cc[r];
                mn = pk_min(mn, v[r]);
            }
        } else {
            unpack_bytes<DPL>(it.pw, v);
            if (PAD) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[r] = pk_max(v[r], padm[r]);
            }
            const uint32_t axis = (uint32_t)(sweep_dir(UP, kind) >> 1);
            const uint32_t p2p2 = __builtin_amdgcn_perm(0u, it.E, 0x0C000C00u | (axis << 16) | axis);
            path_step<DPL, PAD, 16, SLANT>(v, cc, it.pm, it.pm + p2p2, p1p1, padm, mn, edge);
        }
        const uint32_t pm_new = pair_min_line<16>(mn);
        const Packed<DPL> b = pack_bytes<DPL>(v);
        const int slot_w = (j & 1) * ns + sweep_slot(kind, x, x0, tile, band_rows);
        lds_store_bytes<DPL>(lbytes + (size_t)slot_w * dp + k0, b);
        if (g == 0) lmin[slot_w] = pm_new;
        if (j == nrows - 1 && x >= x0 && x < x1) store_packed<DPL, false>(seed_out + kind * wdp + (int64_t)x * dp + k0, b);
    };
    auto step = [&](int kind, int x, int j, uint32_t (&v)[R]) {
        uint32_t cw[CW];
        Item it;
        load_cost(x, j, cw);
        gather(kind, x, j, it);
        finish(kind, x, j, cw, it, v);
    };

    // Band start: the predecessors of row 0 (the previous band's last row) from the seeds into state
    // buffer 1, with their line minima; row 0's costs and penalties staged.
    {
        Stage st;
        stage_load(0, st);
        if (t0 > 0) {
            const int h0 = nrows - 1;
            const int lo1 = max(0, x0 - h0 - 1), hi2 = min(W, x1 + h0 + 1);
            const int n0 = x1 - x0, n1 = (x1 - 1) - lo1, n2 = hi2 - (x0 + 1);
            for (int it = grp; it < n0 + n1 + n2; it += kLinesPerBlock) {
                const int kind = it < n0 ? 0 : (it < n0 + n1 ? 1 : 2);
                const int x = kind == 0 ? x0 + it : (kind == 1 ? lo1 + (it - n0) : x0 + 1 + (it - n0 - n1));
                uint32_t wb[CW], v[R];
                global_load_bytes<DPL>(seed_in + kind * wdp + (int64_t)x * dp + k0, wb);
                unpack_bytes<DPL>(wb, v);
                uint32_t m = 0xFFFFFFFFu;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (PAD) v[r] = pk_max(v[r], padm[r]);
                    m = pk_min(m, v[r]);
                }
                const uint32_t pm = pair_min_line<16>(m);
                const int slot = ns + sweep_slot(kind, x, x0, tile, band_rows);
                lds_store_bytes<DPL>(lbytes + (size_t)slot * dp + k0, pack_bytes<DPL>(v));
                if (g == 0) lmin[slot] = pm;
            }
        }
        stage_store(0, st);
        __syncthreads();
    }

    for (int j = 0; j < nrows; ++j) {
        const int y = row_y(j);
        Stage nx;
        if (j + 1 < nrows) stage_load(j + 1, nx);
        // Core columns x0 + 16 i + grp: all three kinds, summed. Every item's loads first.
        uint32_t cw[kSweepCols][CW];
        Item it[kSweepCols][kKinds];
#pragma unroll
        for (int i = 0; i < kSweepCols; ++i) {
            const int x = x0 + 16 * i + grp;
            if (x < x1) {
                load_cost(x, j, cw[i]);
#pragma unroll
                for (int k = 0; k < kKinds; ++k) gather(k, x, j, it[i][k]);
            }
        }
#pragma unroll
        for (int i = 0; i < kSweepCols; ++i) {
            const int x = x0 + 16 * i + grp;
            if (x < x1) {
                uint32_t s[R], v[R];
                finish(0, x, j, cw[i], it[i][0], s);
                finish(1, x, j, cw[i], it[i][1], v);
#pragma unroll
                for (int r = 0; r < R; ++r) s[r] = pk_add(s[r], v[r]);
                finish(2, x, j, cw[i], it[i][2], v);
#pragma unroll
                for (int r = 0; r < R; ++r) s[r] = pk_add(s[r], v[r]);
                store_words<R>(rec + ((int64_t)y * W + x) * pxs, s);
            }
        }
        // Halo: the left kind on [lo, x0), the right kind on [x1, hi).
        const int lo = row_lo(j), hi = row_hi(j);
        for (int x = x0 - 16 + grp; x >= lo; x -= 16) {
            uint32_t v[R];
            step(1, x, j, v);
        }
        for (int x = x1 + grp; x < hi; x += 16) {
            uint32_t v[R];
            step(2, x, j, v);
        }
        if (j + 1 < nrows) stage_store(j + 1, nx);
        __syncthreads();
    }
}

// Both sweeps of a band in one launch: blockIdx.y 0 walks down, 1 walks up.
template <int DPL, bool PAD, bool SLANT>
__global__ __launch_bounds__(kBlock) void sweep_kernel(LaunchTable t, SgmConsts c, int tile, int band_rows, int band) {
    extern __shared__ uint32_t smem_words[];
    if (blockIdx.y == 0) sweep_body<DPL, PAD, SLANT, false>(t, c, tile, band_rows, band, lds(smem_words));
    else sweep_body<DPL, PAD, SLANT, true>(t, c, tile, band_rows, band, lds(smem_words));
}

}  // namespace

size_t sweep_lds_bytes(int dpl, int tile, int band_rows) {
    const size_t ns = (size_t)sweep_slots(tile, band_rows), nc = (size_t)sweep_cols(tile, band_rows);
    const size_t dp = (size_t)dpl * kGroup;
    return 2 * ns * dp + 2 * ns * 4 + 2 * nc * dp + 2 * 2 * (nc + 1) * 4;
}

bool sweep_shape_ok(int dpl, int tile, int band_rows, size_t lds_limit) {
    if (tile < 16 || tile > 16 * kSweepCols || tile % 16 != 0 || band_rows < 1) return false;
    const size_t nc = (size_t)sweep_cols(tile, band_rows);
    return nc + 1 <= (size_t)kBlock && nc * dpl * kGroup <= (size_t)sweep_stage_loads(dpl) * kBlock * 16 &&
           sweep_lds_bytes(dpl, tile, band_rows) <= lds_limit;
}

int sweep_blocks(int w, int tile) { return (w + tile - 1) / tile; }

hipError_t launch_sweep(int dpl, bool padded, const LaunchTable& t, const SgmConsts& c, int tile, int band_rows, int band,
                        hipStream_t stream) {
    if (t.total_blocks <= 0) return hipSuccess;
    if (!sweep_shape_ok(dpl, tile, band_rows, 160 * 1024) || c.wide_costs || c.ndirs != 8) return hipErrorInvalidValue;
    const size_t lds = sweep_lds_bytes(dpl, tile, band_rows);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(t.total_blocks, 2), dim3(kBlock), lds, stream, t, c, tile, band_rows, band);
    };
    auto pick = [&](auto dplc) {
        constexpr int D = decltype(dplc)::value;
        const int v = (padded ? 1 : 0) | (c.slant_steps ? 2 : 0);
        switch (v) {
            case 0: go(sweep_kernel<D, false, false>); break;
            case 1: go(sweep_kernel<D, true, false>); break;
            case 2: go(sweep_kernel<D, false, true>); break;
            default: go(sweep_kernel<D, true, true>); break;
        }
    };
    switch (dpl) {
        case 2: pick(std::integral_constant<int, 2>{}); break;
        case 4: pick(std::integral_constant<int, 4>{}); break;
        case 8: pick(std::integral_constant<int, 8>{}); break;
        case 12: pick(std::integral_constant<int, 12>{}); break;
        case 16: pick(std::integral_constant<int, 16>{}); break;
        case 24: pick(std::integral_constant<int, 24>{}); break;
        case 32: pick(std::integral_constant<int, 32>{}); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sgm
