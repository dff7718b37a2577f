// Path-kernel instantiations: 32 lanes per scanline (two DPP rows joined), each lane holding half
// the disparities of the 16-lane layout `dpl` (sgm_kernels.h).
#include "sgm_kernels.h"

namespace sgm {

hipError_t launch_paths_g32(int dpl, const PathLaunch& pl, const LaunchTable& t, const SgmConsts& c, hipStream_t stream) {
    switch (dpl) {
        case 4: return launch_paths_variant<2, 32>(pl, t, c, stream);
        case 8: return launch_paths_variant<4, 32>(pl, t, c, stream);
        case 12: return launch_paths_variant<6, 32>(pl, t, c, stream);
        case 16: return launch_paths_variant<8, 32>(pl, t, c, stream);
        case 24: return launch_paths_variant<12, 32>(pl, t, c, stream);
        case 32: return launch_paths_variant<16, 32>(pl, t, c, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace sgm
