// The persistent per-thread schedule, default variant (4 waves per SIMD).
#include "launch.h"
#include "persist_kernel.h"

namespace lam {

void launch_persist(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int n_in, int2 *spill,
                    int occupancy) {
    if (occupancy & kPersistVoteBit)
        launch_persist_vote(grid, block, st, S, F, P, n_in, spill, occupancy & ~kPersistVoteBit);
    else if (occupancy == 0 || occupancy == kPersistWaves)
        k_persist<kPersistWaves><<<grid, block, 0, st>>>(S, F, P, n_in, spill);
    else
        launch_persist_occ(grid, block, st, S, F, P, n_in, spill, occupancy);
}

int persist_blocks_per_cu(int occupancy) {
    if (occupancy & kPersistVoteBit) return persist_vote_blocks_per_cu(occupancy & ~kPersistVoteBit);
    if (occupancy != 0 && occupancy != kPersistWaves) return persist_occ_blocks_per_cu(occupancy);
    int b = 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_persist<kPersistWaves>, kBlock, 0));
    return b;
}

}  // namespace lam
