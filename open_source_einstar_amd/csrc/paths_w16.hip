// Path-kernel instantiations with u16 path costs and step penalties (SgmConsts::wide_costs), 16 lanes
// per scanline, every lane layout (sgm_kernels.h).
#include "sgm_kernels.h"

namespace sgm {

hipError_t launch_paths_w16(int dpl, const PathLaunch& pl, const LaunchTable& t, const SgmConsts& c, hipStream_t stream) {
    switch (dpl) {
        case 2: return launch_paths_variant<2, 16, true>(pl, t, c, stream);
        case 4: return launch_paths_variant<4, 16, true>(pl, t, c, stream);
        case 8: return launch_paths_variant<8, 16, true>(pl, t, c, stream);
        case 12: return launch_paths_variant<12, 16, true>(pl, t, c, stream);
        case 16: return launch_paths_variant<16, 16, true>(pl, t, c, stream);
        case 24: return launch_paths_variant<24, 16, true>(pl, t, c, stream);
        case 32: return launch_paths_variant<32, 16, true>(pl, t, c, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace sgm
