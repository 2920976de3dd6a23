// batch_store.hip -- launcher of the batched gradient (k_batch, multislice.hpp) with stored exit waves in
// every shape and routine, and its occupancy query; the recomputed and checkpointed variants are launched
// from batch_variants.hip.
#include "multislice.hpp"

namespace mset {

hipError_t launch_batch_variant(int n, const PosArgs& A, int grid, int shape, hipStream_t st);

hipError_t launch_batch(int n, const PosArgs& A, int grid, int shape, hipStream_t st) {
    // shape 1: stored waves; shape 2 (1024 threads): the separable 256 routine, stored or checkpointed waves
    if (shape == 1 && A.waves != WAVES_STORE) return hipErrorInvalidValue;
    if (shape == 2 && (n != 256 || !A.sep)) return hipErrorInvalidValue;
    if (shape < 0 || shape > 2) return hipErrorInvalidValue;
    if (A.waves != WAVES_STORE) return launch_batch_variant(n, A, grid, shape, st);
    if (A.sep) {
        if (n != 256 || A.grad_mode != GRAD_REF || A.sset) return hipErrorInvalidValue;
        if (shape == 2) hipLaunchKernelGGL((k_batch<256, 2, WAVES_STORE, true>), dim3(grid), dim3(Cfg<256, 2>::NT), 0, st, A);
        else if (shape) hipLaunchKernelGGL((k_batch<256, 1, WAVES_STORE, true>), dim3(grid), dim3(Cfg<256, 1>::NT), 0, st, A);
        else hipLaunchKernelGGL((k_batch<256, 0, WAVES_STORE, true>), dim3(grid), dim3(Cfg<256>::NT), 0, st, A);
    } else if (shape) {
        MSET_DISPATCH(n, hipLaunchKernelGGL((k_batch<NN, 1, WAVES_STORE, false>), dim3(grid), dim3(Cfg<NN, 1>::NT), 0, st, A))
    } else if (A.pfac && n != 256 && !A.sset) {
        MSET_DISPATCH_SMALL(n, hipLaunchKernelGGL((k_batch<NN, 0, WAVES_STORE, false, true>), dim3(grid), dim3(Cfg<NN>::NT), 0, st, A))
    } else MSET_DISPATCH(n, hipLaunchKernelGGL((k_batch<NN, 0, WAVES_STORE, false>), dim3(grid), dim3(Cfg<NN>::NT), 0, st, A));
    return hipGetLastError();
}

hipError_t occupancy_batch(int n, int shape, int* blocks_per_cu) {
    if (shape)
        MSET_DISPATCH(n, return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (k_batch<NN, 1, WAVES_STORE, false>), Cfg<NN, 1>::NT, 0));
    MSET_DISPATCH(n, return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (k_batch<NN, 0, WAVES_STORE, false>), Cfg<NN>::NT, 0));
    return hipSuccess;
}

}  // namespace mset
