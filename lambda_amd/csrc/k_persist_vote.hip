// The persistent schedule with the wave voting between traversal steps and
// shading (k_persist<W, true>), compiled for 3 and 4 waves per SIMD.
#include "launch.h"
#include "persist_kernel.h"

namespace lam {

void launch_persist_vote(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int n_in, int2 *spill,
                         int occupancy) {
    if (occupancy == 0 || occupancy == 4) k_persist<4, true><<<grid, block, 0, st>>>(S, F, P, n_in, spill);
    else if (occupancy == 3) k_persist<3, true><<<grid, block, 0, st>>>(S, F, P, n_in, spill);
    else throw Failure{LAMBDA_EINVAL, "voting persistent kernel variants: 3 or 4 waves per SIMD"};
}

int persist_vote_blocks_per_cu(int occupancy) {
    int b = 0;
    if (occupancy == 0 || occupancy == 4)
        HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_persist<4, true>, kBlock, 0));
    else if (occupancy == 3) HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_persist<3, true>, kBlock, 0));
    else throw Failure{LAMBDA_EINVAL, "voting persistent kernel variants: 3 or 4 waves per SIMD"};
    return b;
}

}  // namespace lam
