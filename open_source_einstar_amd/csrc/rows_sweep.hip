// Row-kernel instantiations that read band-sweep records (sweep_kernels.hip), every lane layout,
// both row forms (sgm_kernels.h).
#include "sgm_kernels.h"

namespace sgm {

hipError_t launch_rows_sweep(int dpl, bool padded, const LaunchTable& t, const SgmConsts& c, int max_w, hipStream_t stream) {
    if (t.total_blocks <= 0) return hipSuccess;
    if (c.wide_costs || c.ndirs != 8) return hipErrorInvalidValue;
    const bool wide = max_w > kMaxStagedWidth;
    const size_t lds = (size_t)max_w * (wide ? 4 : 15) + 16;
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(t.total_blocks), dim3(kBlock), lds, stream, t, c); };
    auto pick = [&](auto d) {
        constexpr int D = decltype(d)::value;
        auto by_pad = [&](auto pd) {
            constexpr bool PD = decltype(pd)::value;
            if (wide) go(rows_kernel<D, PD, 2, true, false, true>);
            else go(rows_kernel<D, PD, 2, false, false, true>);
        };
        if (padded) by_pad(std::true_type{});
        else by_pad(std::false_type{});
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
