// The sweep kernels that deal a launch's strips upstream first (order 1 of
// strip_of_block), for launches whose strips outnumber the CUs or that share
// the device with another stepping solver: the same sweep bodies as
// k_pcg.hip's, with the other strip order.
#include <atomic>

#include "k_sweep.hpp"

namespace ifl {

template <int KIND, int MODE, int LV, int NG, typename T>
__global__ __launch_bounds__(64 * kWaves) void stream_sweep_up_kernel(StreamArgs a) {
    const int j = strip_of_block(blockIdx.x, a.nlaunch, KIND == SW_BWD, 1);
    if (j < 0) return;
    sweep_body<KIND, MODE, LV, NG, T, false>(a, j);
}

template <int KIND, int MODE, int LV, int NG, typename T>
int launch_sweep_up(Solver* s, const StreamArgs& a, size_t shmem) {
    static std::atomic<bool> attr_set{false};
    if (!attr_set.load(std::memory_order_acquire)) {
        IFL_HIP(hipFuncSetAttribute((const void*)stream_sweep_up_kernel<KIND, MODE, LV, NG, T>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
        attr_set.store(true, std::memory_order_release);
    }
    stream_sweep_up_kernel<KIND, MODE, LV, NG, T><<<sweep_blocks(a.nlaunch, 1), 64 * kWaves, shmem, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// every sweep configuration k_pcg.hip launches
#define IFL_SWEEP_UP(K, M, L, N, T) template int launch_sweep_up<K, M, L, N, T>(Solver*, const StreamArgs&, size_t);
IFL_SWEEP_UP(SW_PRECON, 0, 3, 3, double)
IFL_SWEEP_UP(SW_FWD, 0, 3, 4, double)
IFL_SWEEP_UP(SW_FWD, 0, 4, 4, double)
IFL_SWEEP_UP(SW_FWD, 0, 6, 4, double)
IFL_SWEEP_UP(SW_FWD, 1, 3, 5, double)
IFL_SWEEP_UP(SW_FWD, 1, 4, 5, double)
IFL_SWEEP_UP(SW_FWD, 1, 6, 5, double)
IFL_SWEEP_UP(SW_BWD, 0, 3, 5, double)
IFL_SWEEP_UP(SW_BWD, 0, 4, 6, double)
IFL_SWEEP_UP(SW_BWD, 0, 6, 5, double)
IFL_SWEEP_UP(SW_BWD, 1, 3, 5, double)
IFL_SWEEP_UP(SW_BWD, 1, 4, 6, double)
IFL_SWEEP_UP(SW_BWD, 1, 6, 5, double)
IFL_SWEEP_UP(SW_FWD, 0, 3, 4, float)
IFL_SWEEP_UP(SW_FWD, 0, 4, 4, float)
IFL_SWEEP_UP(SW_FWD, 0, 6, 4, float)
IFL_SWEEP_UP(SW_FWD, 1, 3, 5, float)
IFL_SWEEP_UP(SW_FWD, 1, 4, 5, float)
IFL_SWEEP_UP(SW_FWD, 1, 6, 5, float)
IFL_SWEEP_UP(SW_BWD, 0, 3, 5, float)
IFL_SWEEP_UP(SW_BWD, 0, 4, 6, float)
IFL_SWEEP_UP(SW_BWD, 0, 6, 5, float)
IFL_SWEEP_UP(SW_BWD, 1, 3, 5, float)
IFL_SWEEP_UP(SW_BWD, 1, 4, 6, float)
IFL_SWEEP_UP(SW_BWD, 1, 6, 5, float)
IFL_SWEEP_UP(SW_FWD, 2, 3, 5, float)
IFL_SWEEP_UP(SW_FWD, 2, 4, 5, float)
IFL_SWEEP_UP(SW_FWD, 2, 6, 5, float)
IFL_SWEEP_UP(SW_BWD, 2, 3, 5, float)
IFL_SWEEP_UP(SW_BWD, 2, 4, 6, float)
IFL_SWEEP_UP(SW_BWD, 2, 6, 5, float)
#undef IFL_SWEEP_UP

}  // namespace ifl
