// forward.hip -- launchers of the forward-only kernels (k_fwd, multislice.hpp): one tilt's positions, or the
// positions of several tilts in one persistent launch.
#include "multislice.hpp"

namespace mset {

namespace {
// a separable propagator below 256 is applied as its 1D factors (the general routine's PFAC form)
bool fwd_pfac(int n, const PosArgs& A) { return A.pfac && n != 256 && !A.sset; }
}  // namespace

hipError_t launch_fwd(int n, const PosArgs& A, int grid, hipStream_t st) {
    if (A.sep) {
        if (n != 256 || A.sset) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_fwd<256, false, true>), dim3(grid), dim3(Cfg<256>::NT), 0, st, A);
    } else if (fwd_pfac(n, A)) {
        MSET_DISPATCH_SMALL(n, hipLaunchKernelGGL((k_fwd<NN, false, false, true>), dim3(grid), dim3(Cfg<NN>::NT), 0, st, A))
    } else MSET_DISPATCH(n, hipLaunchKernelGGL((k_fwd<NN, false, false>), dim3(grid), dim3(Cfg<NN>::NT), 0, st, A));
    return hipGetLastError();
}

hipError_t launch_fwd_tilts(int n, const PosArgs& A, int grid, hipStream_t st) {
    if (A.sep) {
        if (n != 256 || A.sset) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_fwd<256, true, true>), dim3(grid), dim3(Cfg<256>::NT), 0, st, A);
    } else if (fwd_pfac(n, A)) {
        MSET_DISPATCH_SMALL(n, hipLaunchKernelGGL((k_fwd<NN, true, false, true>), dim3(grid), dim3(Cfg<NN>::NT), 0, st, A))
    } else MSET_DISPATCH(n, hipLaunchKernelGGL((k_fwd<NN, true, false>), dim3(grid), dim3(Cfg<NN>::NT), 0, st, A));
    return hipGetLastError();
}

}  // namespace mset
