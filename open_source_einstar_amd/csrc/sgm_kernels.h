// Device code shared by the gfx950 SGM kernel translation units: packed-u16 and DPP helpers, the
// launch-table lookup, and the templated path, cost-volume and row kernels. Each .hip file that
// includes it instantiates the variants it launches (sgm_kernels.hip, paths_*.hip, rows_kernels.hip),
// so the variants compile in parallel.
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
#pragma once

#include <algorithm>
#include <climits>
#include <type_traits>

#include "path_lines.h"
#include "sgm_internal.h"

namespace sgm {
namespace {

constexpr float kInvalid = -1.0e9f;
constexpr uint32_t kSentinel = 0x7FFF7FFFu;  // excluded neighbour: 0x7FFF + P1 never wins a min

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t pk_min(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_max(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) + __builtin_bit_cast(u16x2, b));
}
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) - __builtin_bit_cast(u16x2, b));
}
// Saturating at 0xFFFF (v_pk_add_u16 with clamp).
__device__ __forceinline__ uint32_t pk_add_sat(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_add_sat(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}

// DPP controls (gfx9 encoding).
constexpr int kQuadSwap1 = 0xB1;      // quad_perm [1,0,3,2]
constexpr int kQuadSwap2 = 0x4E;      // quad_perm [2,3,0,1]
constexpr int kRowHalfMirror = 0x141;
constexpr int kRowMirror = 0x140;
constexpr int kRowShl1 = 0x101;       // lane i reads lane i+1 of its row
constexpr int kRowShr1 = 0x111;       // lane i reads lane i-1 of its row
constexpr int kWaveShl1 = 0x130;      // lane i reads lane i+1 of the wave
constexpr int kWaveShr1 = 0x138;      // lane i reads lane i-1 of the wave

template <int Ctrl>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, Ctrl, 0xF, 0xF, false);
}
// Neighbour lane's value, or `edge` for the lane at the end of the 16-lane row.
template <int Ctrl>
__device__ __forceinline__ uint32_t dpp_or(uint32_t v, uint32_t edge) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)v, Ctrl, 0xF, 0xF, false);
}

// Neighbour lane's value, or 0 where the shift has no source lane (callers replace those lanes).
template <int Ctrl>
__device__ __forceinline__ uint32_t dpp_zero(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, Ctrl, 0xF, 0xF, true);
}

// Minimum over the 16 lanes of a DPP row, returned to all of them. The DPP moves use the min's
// identity as their old value so the compiler can fold them into v_min_u32_dpp.
template <int Ctrl>
__device__ __forceinline__ uint32_t dpp_id(uint32_t v, uint32_t identity) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)v, Ctrl, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t row_min(uint32_t v) {
    v = min(v, dpp_id<kQuadSwap1>(v, 0xFFFFFFFFu));
    v = min(v, dpp_id<kQuadSwap2>(v, 0xFFFFFFFFu));
    v = min(v, dpp_id<kRowHalfMirror>(v, 0xFFFFFFFFu));
    v = min(v, dpp_id<kRowMirror>(v, 0xFFFFFFFFu));
    return v;
}

__device__ __forceinline__ void locate(const LaunchTable& t, int b, int& pi, int& lb) {
    if (t.n == 1) {
        pi = 0;
        lb = b;
        return;
    }
    int lo = 0, hi = t.n - 1;  // last pair whose first block <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t.first_block[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    pi = lo;
    lb = b - t.first_block[lo];
}

// Descriptors are read through the constant address space (the kernel arguments or the batch's
// device array, both written before the launch), so the block-uniform fields load into scalar
// registers.
#define SGM_CONST __attribute__((address_space(4)))
__device__ __forceinline__ PairDesc load_desc(const LaunchTable& t, int pi) {
    static_assert(sizeof(PairDesc) % 4 == 0, "descriptors are copied as words");
    const SGM_CONST uint32_t* src = t.n == 1 ? (const SGM_CONST uint32_t*)&t.one : (const SGM_CONST uint32_t*)(t.many + pi);
    PairDesc d;
    uint32_t* dst = reinterpret_cast<uint32_t*>(&d);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(PairDesc) / 4); ++i) dst[i] = src[i];
    return d;
}

// LDS pointers carry their address space so that stage bodies compiled out of line still use
// ds_* instructions.
#define SGM_LDS __attribute__((address_space(3)))
template <typename T>
__device__ __forceinline__ SGM_LDS T* lds(T* p) {
    return (SGM_LDS T*)p;
}


// ------------------------------------------------------------------------------------- paths
// aggregate_path, stereo.cpp:44-89, for every start pixel of every direction (:385-409).
// Lane g of a line owns disparities k0..k0+DPL-1 (k0 = g*DPL) as DPL/2 packed u16 pairs
// (low half = even k). Per step:
//   L(k) = C(k) + min(L'(k), L'(k+-1) + P1, [L'(k+-2) + P1], pm + p2) - pm,   pm = min_k L'(k)
// The +-1/+-2 neighbours across lane boundaries come from the adjacent lane by DPP row (or wave)
// shifts, and pm by a 4-step DPP butterfly (joined across two rows by permlane16_swap for 32-lane
// lines).

#define SGM_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ SGM_GLOBAL T* gptr(T* p) {
    return (SGM_GLOBAL T*)p;
}

// DPL path-cost bytes of one pixel, as (DPL+3)/4 words. The loads are non-temporal: each byte is
// read exactly once.
template <int DPL>
__device__ __forceinline__ void load_raw(const SGM_GLOBAL uint8_t* src, uint32_t (&w)[(DPL + 3) / 4]) {
    if constexpr (DPL % 16 == 0) {
#pragma unroll
        for (int i = 0; i < DPL / 16; ++i) {
            const u32x4 q = __builtin_nontemporal_load(reinterpret_cast<const SGM_GLOBAL u32x4*>(src) + i);
            w[4 * i] = q.x;
            w[4 * i + 1] = q.y;
            w[4 * i + 2] = q.z;
            w[4 * i + 3] = q.w;
        }
    } else if constexpr (DPL % 8 == 0) {
#pragma unroll
        for (int i = 0; i < DPL / 8; ++i) {
            const u32x2 q = __builtin_nontemporal_load(reinterpret_cast<const SGM_GLOBAL u32x2*>(src) + i);
            w[2 * i] = q.x;
            w[2 * i + 1] = q.y;
        }
    } else if constexpr (DPL % 4 == 0) {
#pragma unroll
        for (int i = 0; i < DPL / 4; ++i) w[i] = __builtin_nontemporal_load(reinterpret_cast<const SGM_GLOBAL uint32_t*>(src) + i);
    } else {
#pragma unroll
        for (int i = 0; i < DPL / 2; ++i) {
            const uint32_t h = __builtin_nontemporal_load(reinterpret_cast<const SGM_GLOBAL uint16_t*>(src) + i);
            w[i / 2] = (i & 1) ? (w[i / 2] | (h << 16)) : h;
        }
    }
}

// S = the sum over the NDIRS directions of the path-cost bytes (disparities k0..k0+DPL-1), as packed
// u16 pairs. Each word's even and odd bytes are unpacked into u16 lanes by one v_perm each and summed
// with three-input adds: a sum of at most 8 bytes never carries out of its 16-bit lane.
template <int DPL, int NDIRS>
__device__ __forceinline__ void sum_dirs(const uint32_t (&w)[kMaxDirs][(DPL + 3) / 4], uint32_t (&s)[DPL / 2]) {
    static_assert(NDIRS % 2 == 0, "directions come in pairs");
#pragma unroll
    for (int i = 0; i < DPL / 4; ++i) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t sel = h ? 0x0C030C02u : 0x0C010C00u;
            uint32_t acc = __builtin_amdgcn_perm(0u, w[0][i], sel) + __builtin_amdgcn_perm(0u, w[1][i], sel);
#pragma unroll
            for (int d = 2; d < NDIRS; d += 2)
                acc = acc + __builtin_amdgcn_perm(0u, w[d][i], sel) + __builtin_amdgcn_perm(0u, w[d + 1][i], sel);
            s[2 * i + h] = acc;
        }
    }
    if constexpr (DPL % 4 == 2) {
        uint32_t acc = 0;
#pragma unroll
        for (int d = 0; d < NDIRS; ++d) {
            const uint32_t q = w[d][DPL / 4];
            acc += (q & 0xFFu) | ((q & 0xFF00u) << 8);
        }
        s[DPL / 2 - 1] = acc;
    }
}

// Raw buffer views: a wave-uniform base and 32-bit byte offsets. A load at or beyond the view's size
// returns zero instead of touching memory, so a walk can run its prefetch past its last pixel
// without a bounds test per step. (Hosts keep every viewed buffer below 4 GiB.)
typedef __amdgpu_buffer_rsrc_t BufView;
constexpr int kRawBufferWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format, no swizzle
// Base and size are block-uniform; reading them from the first lane keeps the view in scalar
// registers wherever the compiler computed them on the vector unit.
// (readfirstlane returns a signed int: each half goes back to uint32_t before it is widened, or a
// low half with bit 31 set would sign-extend over the high half.)
__device__ __forceinline__ uint32_t first_lane(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ BufView buffer_view(const void* base, uint64_t bytes) {
    const uint32_t n = first_lane(bytes < 0xFFFFFFFFull ? (uint32_t)bytes : 0xFFFFFFFFu);
    const uint64_t a = (uint64_t)base;
    const uint64_t b = ((uint64_t)first_lane((uint32_t)(a >> 32)) << 32) | (uint64_t)first_lane((uint32_t)a);
    return __builtin_amdgcn_make_buffer_rsrc((void*)b, (short)0, (int)n, kRawBufferWord3);
}

// Everything one step of a path reads. Recomputed costs (CV = false): the left census word and the
// lane's DPL right census words (lowest address first, i.e. disparity k0+DPL-1 first). Cost volume
// (CV = true): the lane's DPL cost bytes. Both: the step-penalty word of the pixel whose byte
// holds this step's penalty.
template <int DPL, bool CV>
struct StepIn;
template <int DPL>
struct StepIn<DPL, false> {
    uint64_t a;
    uint64_t w[DPL];
    uint32_t E;  // step penalties of the pixel's axes (census kernel)
};
template <int DPL>
struct StepIn<DPL, true> {
    uint32_t c[(DPL + 3) / 4];
    uint32_t E;
};

// A walk's load positions (byte offsets into the views) and their per-pixel steps.
struct Walk {
    BufView cv, cl, cr, ed;
    uint32_t o_cv, o_cl, o_cr, o_e;
    uint32_t s_cv, s_c, s_e;
};

__device__ __forceinline__ uint64_t u64_of(u32x2 v) { return (uint64_t)v.x | ((uint64_t)v.y << 32); }

// Loads one pixel's inputs at the walk's position and moves the walk to the next pixel.
template <int DPL, bool CV>
__device__ __forceinline__ void load_step(StepIn<DPL, CV>& s, Walk& wk) {
    if constexpr (CV) {
        if constexpr (DPL % 16 == 0) {
#pragma unroll
            for (int i = 0; i < DPL / 16; ++i) {
                const u32x4 q = __builtin_amdgcn_raw_buffer_load_b128(wk.cv, wk.o_cv + 16 * i, 0, 0);
                s.c[4 * i] = q.x;
                s.c[4 * i + 1] = q.y;
                s.c[4 * i + 2] = q.z;
                s.c[4 * i + 3] = q.w;
            }
        } else if constexpr (DPL % 8 == 0) {
#pragma unroll
            for (int i = 0; i < DPL / 8; ++i) {
                const u32x2 q = __builtin_amdgcn_raw_buffer_load_b64(wk.cv, wk.o_cv + 8 * i, 0, 0);
                s.c[2 * i] = q.x;
                s.c[2 * i + 1] = q.y;
            }
        } else if constexpr (DPL % 4 == 0) {
#pragma unroll
            for (int i = 0; i < DPL / 4; ++i) s.c[i] = __builtin_amdgcn_raw_buffer_load_b32(wk.cv, wk.o_cv + 4 * i, 0, 0);
        } else {
#pragma unroll
            for (int i = 0; i < DPL / 2; ++i) {
                const uint32_t h = __builtin_amdgcn_raw_buffer_load_b16(wk.cv, wk.o_cv + 2 * i, 0, 0);
                s.c[i / 2] = (i & 1) ? (s.c[i / 2] | (h << 16)) : h;
            }
        }
        wk.o_cv += wk.s_cv;
    } else {
        s.a = u64_of(__builtin_amdgcn_raw_buffer_load_b64(wk.cl, wk.o_cl, 0, 0));
#pragma unroll
        for (int j = 0; j < DPL; j += 2) {
            const u32x4 q = __builtin_amdgcn_raw_buffer_load_b128(wk.cr, wk.o_cr + 8 * j, 0, 0);
            s.w[j] = (uint64_t)q.x | ((uint64_t)q.y << 32);
            s.w[j + 1] = (uint64_t)q.z | ((uint64_t)q.w << 32);
        }
        wk.o_cl += wk.s_c;
        wk.o_cr += wk.s_c;
    }
    s.E = __builtin_amdgcn_raw_buffer_load_b32(wk.ed, wk.o_e, 0, 0);
    wk.o_e += wk.s_e;
}

// Packed costs of the lane's disparities k0 + 2r, k0 + 2r + 1 (low, high half). Word w[DPL-1-j]
// belongs to disparity k0 + j. With `full` every window word lies in the image (the group-uniform
// common case); otherwise out-of-image words give the invalid cost (stereo.cpp:374-376).
template <int DPL>
__device__ __forceinline__ void census_costs(uint64_t a, const uint64_t (&w)[DPL], bool full, int xr_top, int width,
                                             uint32_t inv, uint32_t (&cc)[DPL / 2]) {
#pragma unroll
    for (int r = 0; r < DPL / 2; ++r) {
        const uint32_t ca = (uint32_t)__popcll(a ^ w[DPL - 1 - 2 * r]);
        const uint32_t cb = (uint32_t)__popcll(a ^ w[DPL - 2 - 2 * r]);
        cc[r] = ca | (cb << 16);
    }
    if (!full) {
#pragma unroll
        for (int r = 0; r < DPL / 2; ++r) {
            const int xa = xr_top - 2 * r, xb = xa - 1;
            const uint32_t lo = (unsigned)xa < (unsigned)width ? (cc[r] & 0xFFFFu) : inv;
            const uint32_t hi = (unsigned)xb < (unsigned)width ? (cc[r] >> 16) : inv;
            cc[r] = lo | (hi << 16);
        }
    }
}

template <int DPL, bool CV>
__device__ __forceinline__ void step_costs(const StepIn<DPL, CV>& s, bool full, int xr_top, int width, uint32_t inv,
                                           uint32_t (&cc)[DPL / 2]) {
    if constexpr (CV) {
        // The cost kernel already wrote the invalid cost outside the image.
#pragma unroll
        for (int i = 0; i < DPL / 4; ++i) {
            cc[2 * i] = __builtin_amdgcn_perm(0u, s.c[i], 0x0C010C00u);
            cc[2 * i + 1] = __builtin_amdgcn_perm(0u, s.c[i], 0x0C030C02u);
        }
        if constexpr (DPL % 4 == 2) {
            const uint32_t q = s.c[DPL / 4];
            cc[DPL / 2 - 1] = (q & 0xFFu) | ((q & 0xFF00u) << 8);
        }
    } else {
        census_costs<DPL>(s.a, s.w, full, xr_top, width, inv, cc);
    }
}

// A pixel's DPL path-cost bytes, packed from the lane's DPL/2 u16 pairs (the low byte of each half).
template <int DPL>
struct Packed {
    uint32_t w[(DPL + 3) / 4];
};
template <int DPL>
__device__ __forceinline__ Packed<DPL> pack_bytes(const uint32_t (&v)[DPL / 2]) {
    Packed<DPL> o;
#pragma unroll
    for (int i = 0; i < DPL / 4; ++i) o.w[i] = __builtin_amdgcn_perm(v[2 * i + 1], v[2 * i], 0x06040200u);
    if constexpr (DPL % 4 == 2) o.w[DPL / 4] = (v[DPL / 2 - 1] & 0xFFu) | ((v[DPL / 2 - 1] >> 8) & 0xFF00u);
    return o;
}
// DPL bytes at dst (aligned to DPL's power-of-two part). Path costs are stored non-temporally
// (NT): they are read back once, by the row kernel, while the cost volume, stored with ordinary
// stores, is re-read by every direction and should stay in the caches.
template <bool NT, typename T, typename P>
__device__ __forceinline__ void put(T v, P* dst) {
    if constexpr (NT) __builtin_nontemporal_store(v, dst);
    else *dst = v;
}
template <int DPL, bool NT = true>
__device__ __forceinline__ void store_packed(SGM_GLOBAL uint8_t* dst, const Packed<DPL>& b) {
    if constexpr (DPL % 16 == 0) {
#pragma unroll
        for (int i = 0; i < DPL / 16; ++i)
            put<NT>(u32x4{b.w[4 * i], b.w[4 * i + 1], b.w[4 * i + 2], b.w[4 * i + 3]},
                    reinterpret_cast<SGM_GLOBAL u32x4*>(dst) + i);
    } else if constexpr (DPL % 8 == 0) {
#pragma unroll
        for (int i = 0; i < DPL / 8; ++i) put<NT>(u32x2{b.w[2 * i], b.w[2 * i + 1]}, reinterpret_cast<SGM_GLOBAL u32x2*>(dst) + i);
    } else if constexpr (DPL % 4 == 0) {
#pragma unroll
        for (int i = 0; i < DPL / 4; ++i) put<NT>(b.w[i], reinterpret_cast<SGM_GLOBAL uint32_t*>(dst) + i);
    } else {
#pragma unroll
        for (int i = 0; i < DPL / 4; ++i) {
            put<NT>((uint16_t)b.w[i], reinterpret_cast<SGM_GLOBAL uint16_t*>(dst) + 2 * i);
            put<NT>((uint16_t)(b.w[i] >> 16), reinterpret_cast<SGM_GLOBAL uint16_t*>(dst) + 2 * i + 1);
        }
        put<NT>((uint16_t)b.w[DPL / 4], reinterpret_cast<SGM_GLOBAL uint16_t*>(dst) + DPL / 2 - 1);
    }
}
template <int DPL, bool NT = true>
__device__ __forceinline__ void store_bytes(SGM_GLOBAL uint8_t* dst, const uint32_t (&v)[DPL / 2]) {
    store_packed<DPL, NT>(dst, pack_bytes<DPL>(v));
}

// A pixel's DPL u16 path costs (W16): the lane's packed pairs as they are.
template <int NW>
struct Words {
    uint32_t w[NW];
};
template <int NW, bool NT = true>
__device__ __forceinline__ void store_words(SGM_GLOBAL uint8_t* dst, const uint32_t (&w)[NW]) {
    if constexpr (NW % 4 == 0) {
#pragma unroll
        for (int i = 0; i < NW / 4; ++i)
            put<NT>(u32x4{w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]}, reinterpret_cast<SGM_GLOBAL u32x4*>(dst) + i);
    } else if constexpr (NW % 2 == 0) {
#pragma unroll
        for (int i = 0; i < NW / 2; ++i) put<NT>(u32x2{w[2 * i], w[2 * i + 1]}, reinterpret_cast<SGM_GLOBAL u32x2*>(dst) + i);
    } else {
#pragma unroll
        for (int i = 0; i < NW; ++i) put<NT>(w[i], reinterpret_cast<SGM_GLOBAL uint32_t*>(dst) + i);
    }
}
// The stored form of a pixel's path costs: bytes (Packed) or u16 (Words).
template <int DPL, bool W16>
using PathOut = std::conditional_t<W16, Words<DPL / 2>, Packed<DPL>>;
template <int DPL, bool W16>
__device__ __forceinline__ PathOut<DPL, W16> make_out(const uint32_t (&v)[DPL / 2]) {
    if constexpr (W16) {
        Words<DPL / 2> o;
#pragma unroll
        for (int i = 0; i < DPL / 2; ++i) o.w[i] = v[i];
        return o;
    } else {
        return pack_bytes<DPL>(v);
    }
}
template <int DPL, bool W16>
__device__ __forceinline__ void store_out(SGM_GLOBAL uint8_t* dst, const PathOut<DPL, W16>& o) {
    if constexpr (W16) store_words<DPL / 2>(dst, o.w);
    else store_packed<DPL>(dst, o);
}

// Edge handling of a 32- or 64-lane line: a wave shift brings the neighbouring line's value (or zero) into
// the line's end lanes, which keep = 0 / sent = sentinel replace. (A select would let the compiler run the
// shift with the end lanes disabled, and a shift never reads a disabled lane's value.)
struct LineEdge {
    uint32_t keep_lo, sent_lo, keep_hi, sent_hi;
};

// One step of the recurrence (stereo.cpp:62-81) on the lane's packed pairs. A line spans G = 16
// lanes (one DPP row), 32 lanes (two rows, exchanging across the row boundary by wave shifts) or 64
// (the whole wave, likewise).
// pmpm holds the previous step's minimum in both halves, jj = pmpm + p2 in both halves.
//
// With slant steps the P1 neighbourhood of k is {k-2, k-1, k+1, k+2}, which is min(a(k-2), a(k+1))
// for the pairwise minima a(j) = min(L'(j), L'(j+1)): one pairwise pass, one shifted pass. An
// excluded neighbour is the sentinel, which never wins a minimum against a real value.
//
// W16 (u16 path costs): the sentinel is 0xFFFF and the P1 candidates are added with saturation; every
// candidate of the minimum is >= the previous step's minimum, and L'(k) <= 0xFFFF is always one of
// them, so a saturated candidate wins exactly when the unsaturated one would (stereo.cpp:71-80 forms
// them in int), and L = C + best - pm then wraps at 16 bits as the reference's u16 cast does.
template <int DPL, bool PAD, int G, bool SLANT, bool W16 = false>
__device__ __forceinline__ void path_step(uint32_t (&prev)[DPL / 2], const uint32_t (&cc)[DPL / 2], uint32_t pmpm,
                                          uint32_t jj, uint32_t p1p1, const uint32_t (&padm)[DPL / 2], uint32_t& mn,
                                          const LineEdge& edge) {
    constexpr int R = DPL / 2;
    constexpr uint32_t kSent = W16 ? 0xFFFFFFFFu : kSentinel;
    uint32_t nb_lo, nb_hi;  // (k0-2, k0-1) and (k0+DPL, k0+DPL+1)
    if constexpr (G == 16) {
        nb_lo = dpp_or<kRowShr1>(prev[R - 1], kSent);
        nb_hi = dpp_or<kRowShl1>(prev[0], kSent);
    } else {
        nb_lo = (dpp_zero<kWaveShr1>(prev[R - 1]) & edge.keep_lo) | edge.sent_lo;
        nb_hi = (dpp_zero<kWaveShl1>(prev[0]) & edge.keep_hi) | edge.sent_hi;
    }
    uint32_t nbr[R];  // packed minimum over the P1 neighbourhood
    if constexpr (SLANT) {
        uint32_t A[R + 2];  // A[r + 1] = (a(k0 + 2r), a(k0 + 2r + 1)), r = -1 .. R
        A[0] = pk_min(nb_lo, __builtin_amdgcn_alignbit(prev[0], nb_lo, 16));
#pragma unroll
        for (int r = 0; r < R; ++r)
            A[r + 1] = pk_min(prev[r], __builtin_amdgcn_alignbit(r + 1 < R ? prev[r + 1] : nb_hi, prev[r], 16));
        const u16x2 h = __builtin_bit_cast(u16x2, nb_hi);
        A[R + 1] = __builtin_bit_cast(uint32_t, __builtin_elementwise_min(h, h.yx));  // low half: a(k0 + DPL)
#pragma unroll
        for (int r = 0; r < R; ++r) nbr[r] = pk_min(A[r], __builtin_amdgcn_alignbit(A[r + 2], A[r + 1], 16));
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t lo = r == 0 ? nb_lo : prev[r - 1];
            const uint32_t hi = r == R - 1 ? nb_hi : prev[r + 1];
            nbr[r] = pk_min(__builtin_amdgcn_alignbit(prev[r], lo, 16), __builtin_amdgcn_alignbit(hi, prev[r], 16));
        }
    }
    mn = 0xFFFFFFFFu;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t cand = W16 ? pk_add_sat(nbr[r], p1p1) : pk_add(nbr[r], p1p1);
        const uint32_t best = pk_min(pk_min(prev[r], cand), jj);
        uint32_t v = pk_add(pk_sub(best, pmpm), cc[r]);
        if (PAD) v = pk_max(v, padm[r]);
        prev[r] = v;
        mn = pk_min(mn, v);
    }
}

// Minimum over the line's lanes of both halves of mn, in both halves of the result; every lane of
// the line receives it. (With equal halves, u32 order is the order of the halves.) Rows pair up as 0-1
// and 2-3 through permlane16_swap (G = 32, 64), then the wave's halves through permlane32_swap (G = 64).
template <int G>
__device__ __forceinline__ uint32_t pair_min_line(uint32_t mn) {
    static_assert(G == 16 || G == 32 || G == 64, "a line spans 16, 32 or 64 lanes");
    const u16x2 h = __builtin_bit_cast(u16x2, mn);
    uint32_t v = row_min(__builtin_bit_cast(uint32_t, __builtin_elementwise_min(h, h.yx)));
    if constexpr (G >= 32) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        v = min((uint32_t)r[0], (uint32_t)r[1]);
    }
    if constexpr (G == 64) {
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        v = min((uint32_t)r[0], (uint32_t)r[1]);
    }
    return v;
}

// Sixteen scanlines (lb * 16 ...) of the pair's path enumeration; no block synchronisation. W16: u16
// path costs and step penalties (SgmConsts::wide_costs).
template <int DPL, bool PAD, bool CV, int G, bool SLANT, bool W16 = false>
__device__ __forceinline__ void paths_body(const PairDesc& P, const SgmConsts& c, int lb) {
    constexpr int R = DPL / 2;
    constexpr uint32_t kSent = W16 ? 0xFFFFFFFFu : kSentinel;
    constexpr uint32_t kPad = W16 ? 0xFFFFu : 0x7FFFu;
    constexpr int ES = W16 ? 2 : 1;  // bytes per stored path cost
    const int line = lb * (int)(blockDim.x / G) + (int)(threadIdx.x / G);  // 4, 8 or 16 lines per block
    if (line >= P.total_lines) return;
    int dir = 0, off = 0;
#pragma unroll
    for (int d = 1; d < kMaxDirs; ++d)
        if (d < P.ndirs && line >= P.line_off[d]) {
            dir = d;
            off = P.line_off[d];
        }
    const int li = line - off;
    const int dx = path_dx(dir), dy = path_dy(dir);
    const int W = P.w, H = P.h;
    const PathLine ln = path_line(dir, li, W, H);

    const int g = threadIdx.x & (G - 1);
    LineEdge edge;
    edge.keep_lo = g == 0 ? 0u : 0xFFFFFFFFu;
    edge.sent_lo = g == 0 ? kSent : 0u;
    edge.keep_hi = g == G - 1 ? 0u : 0xFFFFFFFFu;
    edge.sent_hi = g == G - 1 ? kSent : 0u;
    const int k0 = g * DPL;
    const int min_d = P.min_d;
    const uint32_t inv = (uint32_t)c.invalid_cost;
    const uint32_t p1p1 = (uint32_t)(W16 ? min(c.p1, 0xFFFF) : c.p1) * 0x10001u;
    // The whole group's window [x - min_d - (dp-1), x - min_d] lies in [0, W) for these x.
    const int full_lo = min_d + P.dp - 1, full_hi = W - 1 + min_d;
    uint32_t padm[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        padm[r] = PAD ? ((k0 + 2 * r >= P.nd ? kPad : 0u) | (k0 + 2 * r + 1 >= P.nd ? kPad << 16 : 0u)) : 0u;

    const int x = ln.x, y = ln.y, n = ln.n;
    // Step penalties: byte dir/2 of the word of the pixel stepped into (directions 0, 2, 4, 6 run
    // along their axis) or of the pixel stepped from (1, 3, 5, 7 run against it), selected into
    // both halves.
    const bool along = (dir & 1) == 0;
    const uint32_t axis = (uint32_t)(dir >> 1);
    // W16: half axis & 1 of the pixel's word axis >> 1, in both halves.
    const uint32_t pen_sel = W16 ? ((axis & 1) ? 0x03020302u : 0x01000100u) : (0x0C000C00u | (axis << 16) | axis);
    // Path costs are stored [y][x][dir][dp]: the row kernel then reads one pixel's directions as one
    // contiguous run, and each store here is still a whole dp-byte run.
    const int64_t step_l = ((int64_t)dy * W + dx) * P.lvol_px_stride;
    SGM_GLOBAL uint8_t* lv = gptr(P.lvol) + ((int64_t)y * W + x) * P.lvol_px_stride + ((int64_t)dir * P.dp + k0) * ES;

    // The walk, for one source of matching costs: the u8 cost volume (CW) or recomputation from the
    // census words. A cost-volume launch may still recompute the costs of its first cv_from_dir
    // directions: that trades vector work, which this kernel has to spare, for cost-volume reads,
    // the bandwidth it is bound by.
    auto run = [&](auto cw) {
        constexpr bool CW = decltype(cw)::value;
        // Load positions (the census rows carry pads, so no window word leaves its row).
        Walk wk;
        const int64_t cp = P.census_pitch;
        if constexpr (CW) {
            wk.cv = buffer_view(P.cvol, (uint64_t)W * H * P.dp);
            wk.o_cv = (uint32_t)(((int64_t)y * W + x) * P.dp + k0);
            wk.s_cv = (uint32_t)(((int64_t)dy * W + dx) * P.dp);
        } else {
            wk.cl = buffer_view(P.census_l, (uint64_t)cp * H * 8);
            wk.cr = buffer_view(P.census_r, (uint64_t)cp * H * 8);
            wk.o_cl = (uint32_t)(((int64_t)y * cp + P.census_padl + x) * 8);
            wk.o_cr = (uint32_t)(((int64_t)y * cp + P.census_padl + (x - min_d - k0 - (DPL - 1))) * 8);
            wk.s_c = (uint32_t)(((int64_t)dy * cp + dx) * 8);
        }
        constexpr int kEB = W16 ? 8 : 4;  // penalty bytes per pixel
        wk.ed = buffer_view(P.edge, (uint64_t)W * H * kEB);
        wk.o_e = (uint32_t)((((int64_t)y * W + x) - (along ? 0 : ((int64_t)dy * W + dx))) * kEB + (W16 && axis >= 2 ? 4 : 0));
        wk.s_e = (uint32_t)(((int64_t)dy * W + dx) * kEB);
        // Loads and stores. Two banks of K pixels alternate: while one bank's pixels are computed, the
        // other's loads are in flight. A bank is refilled (2K pixels further on) right after its
        // pixels are computed and before their path costs are stored, so on this hardware, where
        // loads and stores retire through one in-order counter, the wait for a bank's loads never
        // waits for a recent store. Loads past the line's end read other pixels or zeros and are
        // never used.
        // Seven cost-volume pixels a bank: the deepest prefetch that keeps three waves per SIMD at
        // DPL = 8 (160 VGPRs); this walk is bound by the loads in flight, not by occupancy.
        constexpr int K = CW ? (W16 ? 5 : 7) : 2;
        StepIn<DPL, CW> first, bank_a[K], bank_b[K];
        load_step<DPL, CW>(first, wk);  // pixel 0
    #pragma unroll
        for (int j = 0; j < K; ++j) load_step<DPL, CW>(bank_a[j], wk);  // pixels 1 .. K
    #pragma unroll
        for (int j = 0; j < K; ++j) load_step<DPL, CW>(bank_b[j], wk);  // pixels K + 1 .. 2K

        uint32_t prev[R], pmpm;
        {
            uint32_t cc[R];
            step_costs<DPL, CW>(first, x >= full_lo && x <= full_hi, x - min_d - k0, W, inv, cc);
            uint32_t mn = 0xFFFFFFFFu;
    #pragma unroll
            for (int r = 0; r < R; ++r) {
                prev[r] = PAD ? pk_max(cc[r], padm[r]) : cc[r];
                mn = pk_min(mn, prev[r]);
            }
            store_out<DPL, W16>(lv, make_out<DPL, W16>(prev));
            pmpm = pair_min_line<G>(mn);
        }
        // Pixels base .. base + K - 1 from bank bk, then the bank's refill, then the stores. In the main
        // loop every pixel of a half lies on the line (full), so its stores are unconditional and each
        // wait counts exactly the loads and stores issued since; the tail half tests each pixel.
        SGM_GLOBAL uint8_t* lvh = lv + step_l;  // pixel base
        auto half = [&](auto full, StepIn<DPL, CW>(&bk)[K], const int base) {
            constexpr bool kFull = decltype(full)::value;
            // Keeps the scheduler from hoisting the other bank's unpacking (and with it the wait for
            // loads still in flight) into this half.
            __builtin_amdgcn_sched_barrier(0);
            PathOut<DPL, W16> out[K];
    #pragma unroll
            for (int j = 0; j < K; ++j) {
                if (kFull || base + j < n) {
                    const int xc = x + (base + j) * dx;
                    const uint32_t p2p2 = __builtin_amdgcn_perm(0u, bk[j].E, pen_sel);
                    uint32_t cc[R];
                    step_costs<DPL, CW>(bk[j], xc >= full_lo && xc <= full_hi, xc - min_d - k0, W, inv, cc);
                    uint32_t mn;
                    // Byte penalties: pm + p2 < 2^16 in each half; u16: saturated.
                    const uint32_t jj = W16 ? pk_add_sat(pmpm, p2p2) : pmpm + p2p2;
                    path_step<DPL, PAD, G, SLANT, W16>(prev, cc, pmpm, jj, p1p1, padm, mn, edge);
                    out[j] = make_out<DPL, W16>(prev);
                    pmpm = pair_min_line<G>(mn);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
            for (int j = 0; j < K; ++j) load_step<DPL, CW>(bk[j], wk);  // pixels base + 2K ..
    #pragma unroll
            for (int j = 0; j < K; ++j)
                if (kFull || base + j < n) store_out<DPL, W16>(lvh + j * step_l, out[j]);
            lvh += K * step_l;
        };
        int i = 1;
        for (; i + 2 * K <= n; i += 2 * K) {
            half(std::true_type{}, bank_a, i);
            half(std::true_type{}, bank_b, i + K);
        }
        if (i < n) half(std::false_type{}, bank_a, i);
        if (i + K < n) half(std::false_type{}, bank_b, i + K);
    };
    if constexpr (CV) {
        if (dir >= c.cv_from_dir) run(std::true_type{});
        else run(std::false_type{});
    } else {
        run(std::false_type{});
    }
}

template <int DPL, bool PAD, bool CV, int G, bool SLANT, bool W16 = false>
__global__ __launch_bounds__(kBlock) void paths_kernel(LaunchTable t, SgmConsts c) {
    int pi, lb;
    locate(t, blockIdx.x, pi, lb);
    const PairDesc P = load_desc(t, pi);
    paths_body<DPL, PAD, CV, G, SLANT, W16>(P, c, lb);
}

// -------------------------------------------------------------------------------- cost volume
// Matching costs as bytes, [y][x][dp], for the path kernel's cost-volume mode (the cost fill of
// stereo.cpp:366-380, invalid outside the image). A block covers kCostSeg consecutive pixels of a
// row: the right-image census words all of them need (dp + kCostSeg - 1) are staged in LDS once;
// each 16-lane group then takes kCostSeg / 16 consecutive pixels with the path kernel's lane layout.

constexpr int kCostSeg = 128;
// LDS position of window word i: one spare word after every 8, so that the 16 lanes of a group,
// whose windows start 8 words apart, read 16 different bank pairs (a 72-byte lane stride).
__device__ __forceinline__ int win_pos(int i) { return i + (i >> 3); }
constexpr int kWinWords = (512 + kCostSeg) + (512 + kCostSeg) / 8 + 1;

__host__ __device__ inline int cost8_segments(int w) { return (w + kCostSeg - 1) / kCostSeg; }

template <int DPL>
__device__ __forceinline__ void cost8_body(const PairDesc& P, const SgmConsts& c, int lb, SGM_LDS uint64_t* win) {
    const int nseg = cost8_segments(P.w);
    const int y = lb / nseg, x0 = (lb - y * nseg) * kCostSeg;
    const int64_t row = (int64_t)y * P.census_pitch + P.census_padl;
    const int base = x0 - P.min_d - (P.dp - 1);  // xr of win[0]
    const int nwin = P.dp + kCostSeg - 1;
    const int lo = -P.census_padl, hi = P.census_pitch - P.census_padl - 1;  // addressable xr of the row
    for (int i = threadIdx.x; i < nwin; i += kBlock) {
        const int xr = base + i;
        win[win_pos(i)] = (xr >= lo && xr <= hi) ? gptr(P.census_r)[row + xr] : 0ull;
    }
    __syncthreads();
    // Each 16-lane group takes kCostRun consecutive pixels: from one pixel to the next a lane's window
    // moves up by one word, so after the first pixel it reads one new word from LDS instead of DPL.
    constexpr int kCostRun = kCostSeg / kLinesPerBlock;
    const int grp = threadIdx.x >> 4;
    const int k0 = (threadIdx.x & (kGroup - 1)) * DPL;
    const int xs = x0 + grp * kCostRun;
    const int i0 = xs - P.min_d - k0 - (DPL - 1) - base;  // win index of the lane's lowest word at xs
    uint64_t w[DPL];
#pragma unroll
    for (int j = 0; j < DPL; ++j) w[j] = win[win_pos(i0 + j)];
#pragma unroll
    for (int j = 0; j < kCostRun; ++j) {
        if (j > 0) {
#pragma unroll
            for (int t = 0; t + 1 < DPL; ++t) w[t] = w[t + 1];
            w[DPL - 1] = win[win_pos(i0 + j + DPL - 1)];
        }
        const int x = xs + j;
        if (x >= P.w) break;
        const uint64_t a = gptr(P.census_l)[row + x];
        const bool full = x >= P.min_d + P.dp - 1 && x <= P.w - 1 + P.min_d;
        uint32_t cc[DPL / 2];
        census_costs<DPL>(a, w, full, x - P.min_d - k0, P.w, (uint32_t)c.invalid_cost, cc);
        store_bytes<DPL, false>(gptr(P.cvol) + ((int64_t)y * P.w + x) * P.dp + k0, cc);
    }
}

template <int DPL>
__global__ __launch_bounds__(kBlock) void cost8_kernel(LaunchTable t, SgmConsts c) {
    __shared__ uint64_t win[kWinWords];
    int pi, lb;
    locate(t, blockIdx.x, pi, lb);
    const PairDesc P = load_desc(t, pi);
    cost8_body<DPL>(P, c, lb, lds(win));
}

// -------------------------------------------------------------------------------------- rows
// Sum of the path costs, then stereo.cpp:411-461 for one image row per block.

__device__ __forceinline__ float subpixel_offset(float cm, float c0, float cp) {  // stereo.cpp:92-96
    const float denom = cm - 2.0f * c0 + cp;
    if (denom <= 0.0f) return 0.0f;
    const float v = 0.5f * (cm - cp) / denom;
    return v < -0.5f ? -0.5f : (0.5f < v ? 0.5f : v);
}

// Pixels per 16-lane group whose path costs are loaded ahead: three where the registers allow.
constexpr int row_ahead(int dpl) { return dpl <= 8 ? 3 : (dpl <= 16 ? 2 : 1); }
// Threads of the row kernel's large block for a layout: 1024 (at most 128 VGPRs), or 512 where the
// 1024-thread form spills (8 paths with 16 or more disparities per lane, and 32 per lane).
constexpr int big_row_threads(int dpl, int ndirs) { return dpl <= 12 || (ndirs == 4 && dpl <= 24) ? 1024 : 512; }

template <int DPL>
struct RowIn {
    uint32_t w[kMaxDirs][(DPL + 3) / 4];
};
// u16 path costs (W16): the lane's DPL costs of a direction as DPL/2 packed pairs.
template <int DPL>
struct RowIn16 {
    uint32_t w[kMaxDirs][DPL / 2];
};
template <int DPL, bool W16>
using RowInT = std::conditional_t<W16, RowIn16<DPL>, RowIn<DPL>>;
// Band-sweep records (sweep_kernels.hip): the two horizontal directions' path-cost bytes and the
// down and up sweeps' u16 sums of the other six.
template <int DPL>
struct RowInSw {
    uint32_t w[kMaxDirs][(DPL + 3) / 4];  // w[0], w[1]: directions 0 and 1
    uint32_t s[2][DPL / 2];
};

// The direction count is a template parameter so that every load is unconditional: the compiler can
// then wait for each ring slot alone instead of for every load in flight.
template <int DPL, int NDIRS, bool W16>
__device__ __forceinline__ void load_row_in(RowInT<DPL, W16>& in, const SGM_GLOBAL uint8_t* src, int64_t dir_stride) {
#pragma unroll
    for (int d = 0; d < NDIRS; ++d) load_raw<W16 ? 2 * DPL : DPL>(src + d * dir_stride, in.w[d]);
}

// S from u16 path costs: packed 16-bit sums, which wrap as the reference's u16 sum volume does
// (stereo.cpp:82-83).
template <int DPL, int NDIRS>
__device__ __forceinline__ void sum_dirs16(const uint32_t (&w)[kMaxDirs][DPL / 2], uint32_t (&s)[DPL / 2]) {
#pragma unroll
    for (int r = 0; r < DPL / 2; ++r) {
        uint32_t acc = w[0][r];
#pragma unroll
        for (int d = 1; d < NDIRS; ++d) acc = pk_add(acc, w[d][r]);
        s[r] = acc;
    }
}

__device__ __forceinline__ uint32_t half16(uint32_t w, int hi) { return hi ? (w >> 16) : (w & 0xFFFFu); }

// OR over the 16 lanes of a DPP row, in every lane.
__device__ __forceinline__ uint32_t row_or(uint32_t v) {
    v |= dpp_id<kQuadSwap1>(v, 0u);
    v |= dpp_id<kQuadSwap2>(v, 0u);
    v |= dpp_id<kRowHalfMirror>(v, 0u);
    v |= dpp_id<kRowMirror>(v, 0u);
    return v;
}

// S of the lane's element e (disparity k0 + e) from its packed pairs, 0 when e is not in [0, DPL).
template <int DPL>
__device__ __forceinline__ uint32_t lane_elem(const uint32_t (&s)[DPL / 2], int e) {
    uint32_t w = 0;
#pragma unroll
    for (int r = 0; r < DPL / 2; ++r) w = (e >> 1) == r ? s[r] : w;
    w = (e & 1) ? (w >> 16) : (w & 0xFFFFu);
    return (unsigned)e < (unsigned)DPL ? w : 0u;
}

// Image row lb of the pair; every thread of the block calls it (it synchronises twice). smem holds
// 15 bytes per column: the right WTA keys, and the left results that the last pass writes out
// coalesced (keys, second-best sums, disparities, status). WIDE rows (beyond kMaxStagedWidth columns)
// keep only the right WTA keys in LDS (4 bytes per column): each pixel's left results go straight to
// the outputs when computed, and the left/right check reads the disparities back from there.
// SW: the pixels are band-sweep records [S_down[dp] u16, S_up[dp] u16, L0[dp] u8, L1[dp] u8]
// (NDIRS = 2 byte directions).
template <int DPL, bool PAD, int NDIRS, bool WIDE, bool W16 = false, bool SW = false, int RT = kBlock>
__device__ __forceinline__ void rows_body(const PairDesc& P, const SgmConsts& c, int lb, SGM_LDS uint32_t* smem) {
    static_assert(!SW || (NDIRS == 2 && !W16), "sweep records hold two byte directions");
    constexpr int kRowGroups = RT / kGroup;  // 16-lane groups of the block
    using In = std::conditional_t<SW, RowInSw<DPL>, RowInT<DPL, W16>>;
    constexpr int ES = W16 ? 2 : 1;  // bytes per stored path cost
    const int W = P.w, y = lb;
    SGM_LDS uint32_t* rkey = smem;
    SGM_LDS uint32_t* lkey = smem + W;
    SGM_LDS float* drow = (SGM_LDS float*)(smem + 2 * W);
    SGM_LDS uint16_t* lsec = (SGM_LDS uint16_t*)(smem + 3 * W);
    SGM_LDS uint8_t* srow = (SGM_LDS uint8_t*)(smem + 3 * W) + 2 * W;
    for (int x = threadIdx.x; x < W; x += RT) rkey[x] = 0xFFFFFFFFu;
    __syncthreads();

    const int grp = threadIdx.x >> 4, g = threadIdx.x & (kGroup - 1);
    const int k0 = g * DPL;
    const int nd = P.nd, min_d = P.min_d;
    const int64_t dp = P.dp * ES;  // bytes between a pixel's directions
    const int64_t pxs = P.lvol_px_stride;  // [y][x][dir][dp]: a pixel's directions are contiguous
    const SGM_GLOBAL uint8_t* lrow = gptr(P.lvol) + (int64_t)y * W * pxs + (SW ? 4 * P.dp : 0) + k0 * ES;
    const SGM_GLOBAL uint8_t* sums = gptr(P.lvol) + (int64_t)y * W * pxs + 2 * k0;  // SW: the u16 sums
    const int64_t row0 = (int64_t)y * W;
    auto load_in = [&](In& in, int x) {
        if constexpr (SW) {
            load_raw<DPL>(lrow + (int64_t)x * pxs, in.w[0]);
            load_raw<DPL>(lrow + (int64_t)x * pxs + dp, in.w[1]);
            load_raw<2 * DPL>(sums + (int64_t)x * pxs, in.s[0]);
            load_raw<2 * DPL>(sums + (int64_t)x * pxs + 2 * P.dp, in.s[1]);
        } else {
            load_row_in<DPL, NDIRS, W16>(in, lrow + (int64_t)x * pxs, dp);
        }
    };

    // One pixel's results between its computation and its commit (right-WTA atomics and stores).
    struct PxOut {
        uint32_t s[DPL / 2];  // S of the lane's disparities, packed pairs
        uint32_t key, sec;    // best (S << 16 | k) over the row's lanes, second best S
        float d;
        uint8_t st;
    };
    // Sums, left WTA with uniqueness and subpixel (stereo.cpp:412-432); no memory side effects, so
    // the computations of several pixels can overlap.
    auto compute = [&](const In& in, int x, PxOut& o) {
        if constexpr (SW) {
            sum_dirs<DPL, 2>(in.w, o.s);
#pragma unroll
            for (int r = 0; r < DPL / 2; ++r) o.s[r] = pk_add(o.s[r], pk_add(in.s[0][r], in.s[1][r]));
        } else if constexpr (W16) {
            sum_dirs16<DPL, NDIRS>(in.w, o.s);
        } else {
            sum_dirs<DPL, NDIRS>(in.w, o.s);
        }

        // stereo.cpp:419-421: candidates whose right pixel x - (min_d + k) lies in the image.
        const int min_k = max(0, x - min_d - (W - 1));
        const int max_k = min(nd, x - min_d + 1);
        uint32_t key = 0xFFFFFFFFu;
        uint32_t sec = 0xFFFFu;
        if (!PAD && min_k == 0 && max_k == nd) {
            // Every disparity is a candidate (most pixels): no per-element window tests.
#pragma unroll
            for (int j = 0; j < DPL; ++j)
                key = min(key, ((j & 1) ? (o.s[j >> 1] & 0xFFFF0000u) : (o.s[j >> 1] << 16)) | (uint32_t)(k0 + j));
            key = row_min(key);
            const int bk = (int)(key & 0xFFFFu);
            // Second best over |k - bk| > 1: the excluded halves are forced to 0xFFFF. Bit e of xm marks
            // element e (disparity k0 + e) as one of bk - 1, bk, bk + 1.
            // (The shift reaches DPL + 3 bits: a 64-bit word where that passes 31.)
            using Mask = std::conditional_t<(DPL + 3 < 32), uint32_t, uint64_t>;
            const int t = min(max(bk - k0, -2), DPL + 1);
            const Mask xm = ((Mask)7 << (t + 2)) >> 3;
            uint32_t m = 0xFFFFFFFFu;
#pragma unroll
            for (int r = 0; r < DPL / 2; ++r) {
                const uint32_t lo = 0u - (uint32_t)((xm >> (2 * r)) & 1u), hi = 0u - (uint32_t)((xm >> (2 * r + 1)) & 1u);
                m = pk_min(m, o.s[r] | (lo & 0xFFFFu) | (hi & 0xFFFF0000u));
            }
            sec = min(m & 0xFFFFu, m >> 16);
        } else {
#pragma unroll
            for (int j = 0; j < DPL; ++j) {
                const int k = k0 + j;
                if (k >= min_k && k < max_k) key = min(key, (half16(o.s[j >> 1], j & 1) << 16) | (uint32_t)k);
            }
            key = row_min(key);  // lowest S, then lowest k: the first strict argmin of :422-424
            const int bk0 = key == 0xFFFFFFFFu ? -1 : (int)(key & 0xFFFFu);
#pragma unroll
            for (int j = 0; j < DPL; ++j) {  // stereo.cpp:425-427
                const int k = k0 + j;
                const int dk = k - bk0;
                if (k >= min_k && k < max_k && (dk > 1 || dk < -1)) sec = min(sec, half16(o.s[j >> 1], j & 1));
            }
        }
        const int bk = key == 0xFFFFFFFFu ? -1 : (int)(key & 0xFFFFu);
        const uint32_t bS = key >> 16;
        sec = row_min(sec);
        float d = kInvalid;
        uint8_t st = 0;
        if (bk >= 0) {
            st = 1;
            if (!(sec != 0xFFFFu && (float)bS >= c.uniqueness * (float)sec)) {  // stereo.cpp:428
                st |= 2;
                d = (float)(min_d + bk);
                if (c.subpixel && bk > min_k && bk + 1 < max_k) {  // stereo.cpp:430-431
                    // S(bk-1) and S(bk+1) in one word: the owning lanes contribute their halves, the
                    // others 0, and one OR over the row gathers both.
                    const uint32_t vv = row_or((lane_elem<DPL>(o.s, bk - 1 - k0) << 16) | lane_elem<DPL>(o.s, bk + 1 - k0));
                    d += subpixel_offset((float)(vv >> 16), (float)bS, (float)(vv & 0xFFFFu));
                }
            }
        }
        o.key = key;
        o.sec = sec;
        o.d = d;
        o.st = st;
    };
    // Right WTA by LDS atomic minimum of the same keys (stereo.cpp:434-447), then the pixel's
    // outputs.
    auto commit = [&](const PxOut& o, int x) {
        const int min_k = max(0, x - min_d - (W - 1));
        const int max_k = min(nd, x - min_d + 1);
        if (!PAD && min_k == 0 && max_k == nd) {
            SGM_LDS uint32_t* rk = rkey + (x - min_d - k0 - (DPL - 1));  // rk[DPL-1-j]: right pixel of k0+j
#pragma unroll
            for (int j = 0; j < DPL; ++j) {
                const uint32_t kk = ((j & 1) ? (o.s[j >> 1] & 0xFFFF0000u) : (o.s[j >> 1] << 16)) | (uint32_t)(k0 + j);
                __hip_atomic_fetch_min(&rk[DPL - 1 - j], kk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        } else {
#pragma unroll
            for (int j = 0; j < DPL; ++j) {
                const int k = k0 + j;
                if (k >= min_k && k < max_k)
                    __hip_atomic_fetch_min(&rkey[x - min_d - k], (half16(o.s[j >> 1], j & 1) << 16) | (uint32_t)k,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        if (g == 0) {
            if constexpr (WIDE) {
                const int bk = o.key == 0xFFFFFFFFu ? -1 : (int)(o.key & 0xFFFFu);
                P.disp[row0 + x] = o.d;  // before the left/right check, which reads it back
                if (P.status) P.status[row0 + x] = o.st;
                if (P.dint) P.dint[row0 + x] = bk >= 0 ? min_d + bk : INT_MIN;
                if (P.best) P.best[row0 + x] = (uint16_t)(bk >= 0 ? o.key >> 16 : 0xFFFFu);
                if (P.second) P.second[row0 + x] = (uint16_t)(bk >= 0 ? o.sec : 0xFFFFu);
            } else {
                drow[x] = o.d;
                srow[x] = o.st;
                lkey[x] = o.key;
                lsec[x] = (uint16_t)o.sec;
            }
        }
        if (P.sum_volume) {
            uint16_t* sv = P.sum_volume + (row0 + x) * nd;
#pragma unroll
            for (int j = 0; j < DPL; ++j)
                if (k0 + j < nd) sv[k0 + j] = (uint16_t)half16(o.s[j >> 1], j & 1);
        }
    };

    // Pixels grp, grp + kRowGroups, ...: a ring of kRowAhead pixels' path costs is in flight while the
    // previous ring is computed; the ring's pixels are computed together (their work overlaps),
    // the slots refilled, then the pixels committed. Reads past the row's end are clamped to its
    // last pixel and never used.
    constexpr int kRowAhead = row_ahead(W16 ? 2 * DPL : DPL);
    In ring[kRowAhead];
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) load_in(ring[j], min(grp + j * kRowGroups, W - 1));
    for (int x = grp; x < W; x += kRowAhead * kRowGroups) {
        PxOut o[kRowAhead];
#pragma unroll
        for (int j = 0; j < kRowAhead; ++j) {
            const int xj = x + j * kRowGroups;
            if (xj < W) compute(ring[j], xj, o[j]);
        }
#pragma unroll
        for (int j = 0; j < kRowAhead; ++j) load_in(ring[j], min(x + (j + kRowAhead) * kRowGroups, W - 1));
#pragma unroll
        for (int j = 0; j < kRowAhead; ++j) {
            const int xj = x + j * kRowGroups;
            if (xj < W) commit(o[j], xj);
        }
    }
    __syncthreads();
    // Left/right consistency, stereo.cpp:450-461 (lround: halves away from zero).
    for (int x = threadIdx.x; x < W; x += RT) {
        const uint32_t rk = rkey[x];
        float d;
        uint8_t st = 0;
        if constexpr (WIDE) {
            d = P.disp[row0 + x];  // written above by this block (the barrier orders it)
        } else {
            d = drow[x];
            st = srow[x];
        }
        if (d > -1.0e8f) {
            const long xr = lroundf((float)x - d);
            bool ok = xr >= 0 && xr < W;
            if (ok) {
                const uint32_t k2 = rkey[xr];
                ok = k2 != 0xFFFFFFFFu && !(fabsf((float)(min_d + (int)(k2 & 0xFFFFu)) - d) > (float)c.lr_max_diff);
            }
            if (ok) st |= 4;
            else d = kInvalid;
        }
        P.disp[row0 + x] = d;
        if constexpr (WIDE) {
            if (P.status && (st & 4)) P.status[row0 + x] |= 4;
        } else {
            if (P.status) P.status[row0 + x] = st;
            const uint32_t lk = lkey[x];
            const int bk = lk == 0xFFFFFFFFu ? -1 : (int)(lk & 0xFFFFu);
            if (P.dint) P.dint[row0 + x] = bk >= 0 ? min_d + bk : INT_MIN;
            if (P.best) P.best[row0 + x] = (uint16_t)(bk >= 0 ? lk >> 16 : 0xFFFFu);
            if (P.second) P.second[row0 + x] = (uint16_t)(bk >= 0 ? lsec[x] : 0xFFFFu);
        }
        if (P.rdisp) P.rdisp[row0 + x] = rk == 0xFFFFFFFFu ? INT_MIN : min_d + (int)(rk & 0xFFFFu);
        if (P.conf) P.conf[row0 + x] = d > -1.0e8f ? 1.0f : 0.0f;
    }
}

// RT threads per block, one image row a block: 256, or 1024 for pairs of few rows (four times the
// waves on a CU and a quarter of each group's pixels).
template <int DPL, bool PAD, int NDIRS, bool WIDE, bool W16 = false, bool SW = false, int RT = kBlock>
__global__ __launch_bounds__(RT) void rows_kernel(LaunchTable t, SgmConsts c) {
    extern __shared__ uint32_t smem[];
    int pi, lb;
    locate(t, blockIdx.x, pi, lb);
    const PairDesc P = load_desc(t, pi);
    rows_body<DPL, PAD, NDIRS, WIDE, W16, SW, RT>(P, c, lb, lds(smem));
}


}  // namespace

// Path-kernel variant selection shared by the translation units that instantiate it.
struct PathLaunch {
    bool padded, cv, slant;
    unsigned threads;
};

namespace {

template <int D, int G, bool W16 = false>
hipError_t launch_paths_variant(const PathLaunch& pl, const LaunchTable& t, const SgmConsts& c, hipStream_t stream) {
    const dim3 grid(t.total_blocks), block(pl.threads);
    auto go = [&](auto pad, auto cv) {
        constexpr bool PD = decltype(pad)::value, CV = decltype(cv)::value;
        if (pl.slant) hipLaunchKernelGGL((paths_kernel<D, PD, CV, G, true, W16>), grid, block, 0, stream, t, c);
        else hipLaunchKernelGGL((paths_kernel<D, PD, CV, G, false, W16>), grid, block, 0, stream, t, c);
    };
    if (pl.cv) {
        if (pl.padded) go(std::true_type{}, std::true_type{});
        else go(std::false_type{}, std::true_type{});
    } else {
        if (pl.padded) go(std::true_type{}, std::false_type{});
        else go(std::false_type{}, std::false_type{});
    }
    return hipGetLastError();
}

}  // namespace

// Path-kernel launchers by lane layout, one translation unit each: 16 lanes per scanline with
// dpl <= 8 (paths_g16_lo.hip) or dpl >= 12 (paths_g16_hi.hip), 32 lanes (paths_g32.hip; dpl is
// the 16-lane layout's, each lane holding half of it) and 64 lanes (paths_g64.hip; a quarter).
hipError_t launch_paths_g16_lo(int dpl, const PathLaunch& pl, const LaunchTable& t, const SgmConsts& c, hipStream_t stream);
hipError_t launch_paths_g16_hi(int dpl, const PathLaunch& pl, const LaunchTable& t, const SgmConsts& c, hipStream_t stream);
hipError_t launch_paths_g32(int dpl, const PathLaunch& pl, const LaunchTable& t, const SgmConsts& c, hipStream_t stream);
hipError_t launch_paths_g64(int dpl, const PathLaunch& pl, const LaunchTable& t, const SgmConsts& c, hipStream_t stream);
// u16 path costs (SgmConsts::wide_costs), 16 lanes per scanline (paths_w16.hip), and their row kernels
// (rows_w16.hip).
hipError_t launch_paths_w16(int dpl, const PathLaunch& pl, const LaunchTable& t, const SgmConsts& c, hipStream_t stream);
hipError_t launch_rows_w16(int dpl, bool padded, const LaunchTable& t, const SgmConsts& c, int max_w, hipStream_t stream);

}  // namespace sgm
