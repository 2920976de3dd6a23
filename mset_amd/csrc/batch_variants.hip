// batch_variants.hip -- launcher of the batched gradient (k_batch, multislice.hpp) with recomputed or
// checkpointed exit waves (the default workgroup shape; for the separable 256 checkpoint routine also the
// short-launch shape).
#include "multislice.hpp"

namespace mset {

hipError_t launch_batch_variant(int n, const PosArgs& A, int grid, int shape, hipStream_t st) {
    if (shape != 0 && (shape != 2 || A.waves != WAVES_CHECKPOINT || !A.sep)) return hipErrorInvalidValue;
    if (A.waves == WAVES_RECOMPUTE) {
        if (A.sep) return hipErrorInvalidValue;
        MSET_DISPATCH(n, hipLaunchKernelGGL((k_batch<NN, 0, WAVES_RECOMPUTE, false>), dim3(grid), dim3(Cfg<NN>::NT), 0, st, A))
    } else if (A.waves == WAVES_CHECKPOINT) {
        if (A.sset || A.ckseg < 2) return hipErrorInvalidValue;
        if (A.sep) {
            if (n != 256 || A.grad_mode != GRAD_REF) return hipErrorInvalidValue;
            if (shape == 2)
                hipLaunchKernelGGL((k_batch<256, 2, WAVES_CHECKPOINT, true>), dim3(grid), dim3(Cfg<256, 2>::NT), 0, st, A);
            else
                hipLaunchKernelGGL((k_batch<256, 0, WAVES_CHECKPOINT, true>), dim3(grid), dim3(Cfg<256>::NT), 0, st, A);
        } else {
            MSET_DISPATCH(n, hipLaunchKernelGGL((k_batch<NN, 0, WAVES_CHECKPOINT, false>), dim3(grid), dim3(Cfg<NN>::NT), 0, st, A))
        }
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mset
