// Row-kernel instantiations that sum u16 path costs (SgmConsts::wide_costs), every lane layout, both
// row forms (sgm_kernels.h).
#include "sgm_kernels.h"

namespace sgm {

hipError_t launch_rows_w16(int dpl, bool padded, const LaunchTable& t, const SgmConsts& c, int max_w, hipStream_t stream) {
    const bool wide = max_w > kMaxStagedWidth;
    const size_t lds = (size_t)max_w * (wide ? 4 : 15) + 16;
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(t.total_blocks), dim3(kBlock), lds, stream, t, c); };
    auto pick = [&](auto d) {
        constexpr int D = decltype(d)::value;
        auto by_dirs = [&](auto pd) {
            constexpr bool PD = decltype(pd)::value;
            if (c.ndirs == 8) {
                if (wide) go(rows_kernel<D, PD, 8, true, true>);
                else go(rows_kernel<D, PD, 8, false, true>);
            } else {
                if (wide) go(rows_kernel<D, PD, 4, true, true>);
                else go(rows_kernel<D, PD, 4, false, true>);
            }
        };
        if (padded) by_dirs(std::true_type{});
        else by_dirs(std::false_type{});
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
