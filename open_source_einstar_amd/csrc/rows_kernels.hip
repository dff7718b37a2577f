// Row-kernel and byte cost-volume instantiations for every lane layout (sgm_kernels.h).
#include "sgm_kernels.h"

namespace sgm {

#define SGM_FOR_DPL(X) X(2) X(4) X(8) X(12) X(16) X(24) X(32)

hipError_t launch_cost8(int dpl, const LaunchTable& t, const SgmConsts& c, hipStream_t stream) {
    if (t.total_blocks <= 0) return hipSuccess;
    switch (dpl) {
#define SGM_CASE(D)                                                                                     \
    case D:                                                                                             \
        hipLaunchKernelGGL((cost8_kernel<D>), dim3(t.total_blocks), dim3(kBlock), 0, stream, t, c);     \
        break;
        SGM_FOR_DPL(SGM_CASE)
#undef SGM_CASE
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rows(int dpl, bool padded, const LaunchTable& t, const SgmConsts& c, int max_w, hipStream_t stream,
                       int threads) {
    if (t.total_blocks <= 0) return hipSuccess;
    if (c.ndirs != 4 && c.ndirs != 8) return hipErrorInvalidValue;
    if (c.wide_costs) return launch_rows_w16(dpl, padded, t, c, max_w, stream);
    // Rows up to kMaxStagedWidth stage their left results in LDS (15 bytes per column); wider rows keep
    // only the right WTA keys there (4 bytes per column).
    const bool wide = max_w > kMaxStagedWidth;
    const size_t lds = (size_t)max_w * (wide ? 4 : 15) + 16;
    const bool big = threads == 1024 && !wide;
    auto go = [&](auto kern, int nt) { hipLaunchKernelGGL(kern, dim3(t.total_blocks), dim3(nt), lds, stream, t, c); };
    switch (dpl) {
#define SGM_ROWS_N(D, PD, N)                                                                      \
    do {                                                                                          \
        if (wide) go(rows_kernel<D, PD, N, true>, kBlock);                                        \
        else if (big) go(rows_kernel<D, PD, N, false, false, false, big_row_threads(D, N)>,      \
                         big_row_threads(D, N));                                                  \
        else go(rows_kernel<D, PD, N, false>, kBlock);                                            \
    } while (0)
#define SGM_ROWS(D, PD)                                   \
    do {                                                  \
        if (c.ndirs == 8) SGM_ROWS_N(D, PD, 8);           \
        else SGM_ROWS_N(D, PD, 4);                        \
    } while (0)
#define SGM_CASE(D)                  \
    case D:                          \
        if (padded) SGM_ROWS(D, true); \
        else SGM_ROWS(D, false);     \
        break;
        SGM_FOR_DPL(SGM_CASE)
#undef SGM_CASE
#undef SGM_ROWS
#undef SGM_ROWS_N
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sgm
