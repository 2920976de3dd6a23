// multislice.hip -- launchers of the sequential and plane-FFT kernels (multislice.hpp).
#include "multislice.hpp"

namespace mset {

bool plane_size_supported(int n) {
    return n == 16 || n == 24 || n == 32 || n == 48 || n == 64 || n == 96 || n == 128 || n == 256;
}

bool plane_in_global(int n) { return n == 256; }

int position_threads(int n) {
    switch (n) {
        case 16: return Cfg<16>::NT;
        case 24: return Cfg<24>::NT;
        case 32: return Cfg<32>::NT;
        case 48: return Cfg<48>::NT;
        case 64: return Cfg<64>::NT;
        case 96: return Cfg<96>::NT;
        case 128: return Cfg<128>::NT;
        case 256: return Cfg<256>::NT;
        default: return 0;
    }
}

hipError_t launch_plain(int n, const PosArgs& A, int k, hipStream_t st) {
    if (A.pfac && n != 256 && !A.sset)
        MSET_DISPATCH_SMALL(n, hipLaunchKernelGGL((k_plain<NN, true>), dim3(1), dim3(Cfg<NN>::NT), 0, st, A, k))
    else
        MSET_DISPATCH(n, hipLaunchKernelGGL((k_plain<NN, false>), dim3(1), dim3(Cfg<NN>::NT), 0, st, A, k));
    return hipGetLastError();
}


hipError_t launch_plane_fft(int n, const float2* in, float2* out, const float2* tw, int dir, float alpha, float beta,
                            float2* planes, hipStream_t st, float2* out2, float alpha2, float beta2) {
    if (n == 256) {
        // rows of `in` into the scratch plane, then its columns into `out`: 2 x 64 workgroups
        constexpr int N = 256, TL = 4, NT = 64;
        if (!planes) return hipErrorInvalidValue;
        if (dir == 0) {
            hipLaunchKernelGGL((k_plane_lines<N, TL, NT, false, true>), dim3(N / TL), dim3(NT), 0, st, in, planes, tw,
                               1.f, 0.f, nullptr, 0.f, 0.f);
            hipLaunchKernelGGL((k_plane_lines<N, TL, NT, false, false>), dim3(N / TL), dim3(NT), 0, st, planes, out, tw,
                               1.f, 0.f, nullptr, 0.f, 0.f);
        } else {
            hipLaunchKernelGGL((k_plane_lines<N, TL, NT, true, true>), dim3(N / TL), dim3(NT), 0, st, in, planes, tw,
                               1.f, 0.f, nullptr, 0.f, 0.f);
            hipLaunchKernelGGL((k_plane_lines<N, TL, NT, true, false>), dim3(N / TL), dim3(NT), 0, st, planes, out, tw,
                               alpha, beta, out2, alpha2, beta2);
        }
        return hipGetLastError();
    }
    // smaller planes: one workgroup transforms the whole plane on chip
    MSET_DISPATCH_SMALL(n, hipLaunchKernelGGL(k_plane_fft<NN>, dim3(1), dim3(Cfg<NN>::NT), 0, st, in, out, tw, dir, alpha,
                                              beta, planes, out2, alpha2, beta2));
    return hipGetLastError();
}

}  // namespace mset
