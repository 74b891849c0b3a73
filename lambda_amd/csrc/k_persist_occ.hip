// The persistent per-thread schedule compiled for 2 and 3 waves per SIMD
// (autotuner candidates beside the default 4).
#include "launch.h"
#include "persist_kernel.h"

namespace lam {

void launch_persist_occ(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int n_in, int2 *spill,
                        int occupancy) {
    if (occupancy == 2) k_persist<2><<<grid, block, 0, st>>>(S, F, P, n_in, spill);
    else if (occupancy == 3) k_persist<3><<<grid, block, 0, st>>>(S, F, P, n_in, spill);
    else throw Failure{LAMBDA_EINVAL, "persistent kernel variants: 2, 3 or 4 waves per SIMD"};
}

int persist_occ_blocks_per_cu(int occupancy) {
    int b = 0;
    if (occupancy == 2) HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_persist<2>, kBlock, 0));
    else if (occupancy == 3) HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_persist<3>, kBlock, 0));
    else throw Failure{LAMBDA_EINVAL, "persistent kernel variants: 2, 3 or 4 waves per SIMD"};
    return b;
}

}  // namespace lam
