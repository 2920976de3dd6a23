// tv.hip -- the total-variation regulariser on the device (TV_method.m:5-84, 3D).
//
// Chambolle's projection for min_u 1/2 |u - f|^2 + lambda TV(u) with periodic forward differences:
//   D = TV_div(P) - f / lambda                      (k_tv_dual_residual)
//   P = TV_project(P + tau TV_del(D)), tau = 1/12   (k_tv_project)
// repeated n_iter times from P = 0, then u = f - lambda TV_div(P) and the error lambda sum |TV_del(u)|.
// The dual field P is three float volumes (structure of arrays, so every stencil read is coalesced);
// the operation order of each element follows TV_method.m so that single-precision results match a
// MATLAB single run to rounding.  Volumes are [z][y][x] with MATLAB dims (x, y, z) = (1, 2, 3):
// circshift(A, [1 0 0]) is A(x - 1) and circshift(A, [-1 0 0]) is A(x + 1), periodic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "kernels.hpp"

namespace mset {

namespace {

struct Grid3 {
    int nx, ny, nz;
    __device__ __forceinline__ void coords(long long i, int& x, int& y, int& z) const {
        x = (int)(i % nx);
        const long long r = i / nx;
        y = (int)(r % ny);
        z = (int)(r / ny);
    }
    __device__ __forceinline__ long long at(int x, int y, int z) const { return ((long long)z * ny + y) * nx + x; }
};

// TV_div: P1 + P2 + P3 - P1(x-1) - P2(y-1) - P3(z-1)
__device__ __forceinline__ float tv_div(const Grid3& g, const float* P0, const float* P1, const float* P2, int x, int y,
                                        int z) {
    const long long i = g.at(x, y, z);
    float d = P0[i] + P1[i] + P2[i];
    d = d - P0[g.at(x == 0 ? g.nx - 1 : x - 1, y, z)];
    d = d - P1[g.at(x, y == 0 ? g.ny - 1 : y - 1, z)];
    d = d - P2[g.at(x, y, z == 0 ? g.nz - 1 : z - 1)];
    return d;
}

__global__ void k_tv_dual_residual(Grid3 g, const float* __restrict__ P0, const float* __restrict__ P1,
                                   const float* __restrict__ P2, const float* __restrict__ f, float lambda,
                                   float* __restrict__ D) {
    const long long nv = (long long)g.nx * g.ny * g.nz;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nv; i += (long long)gridDim.x * blockDim.x) {
        int x, y, z;
        g.coords(i, x, y, z);
        D[i] = tv_div(g, P0, P1, P2, x, y, z) - f[i] / lambda;
    }
}

__global__ void k_tv_project(Grid3 g, float* __restrict__ P0, float* __restrict__ P1, float* __restrict__ P2,
                             const float* __restrict__ D) {
    const float tau = 1.0f / 12.0f;
    const long long nv = (long long)g.nx * g.ny * g.nz;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nv; i += (long long)gridDim.x * blockDim.x) {
        int x, y, z;
        g.coords(i, x, y, z);
        const float d = D[i];
        const float q0 = P0[i] + tau * (D[g.at(x + 1 == g.nx ? 0 : x + 1, y, z)] - d);
        const float q1 = P1[i] + tau * (D[g.at(x, y + 1 == g.ny ? 0 : y + 1, z)] - d);
        const float q2 = P2[i] + tau * (D[g.at(x, y, z + 1 == g.nz ? 0 : z + 1)] - d);
        const float den = sqrtf(fmaxf(1.0f, q0 * q0 + q1 * q1 + q2 * q2));
        P0[i] = q0 / den;
        P1[i] = q1 / den;
        P2[i] = q2 / den;
    }
}

__global__ void k_tv_output(Grid3 g, const float* __restrict__ P0, const float* __restrict__ P1,
                            const float* __restrict__ P2, const float* f, float lambda, float* u) {
    const long long nv = (long long)g.nx * g.ny * g.nz;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nv; i += (long long)gridDim.x * blockDim.x) {
        int x, y, z;
        g.coords(i, x, y, z);
        u[i] = f[i] - lambda * tv_div(g, P0, P1, P2, x, y, z);   // f and u may alias (same element)
    }
}

constexpr int RED_THREADS = 256;

// per-block partial sums (double) of |TV_del(u)|, summed on the host in block order
__global__ void k_tv_error(Grid3 g, const float* __restrict__ u, double* __restrict__ part) {
    __shared__ double red[RED_THREADS];
    const long long nv = (long long)g.nx * g.ny * g.nz;
    double s = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nv; i += (long long)gridDim.x * blockDim.x) {
        int x, y, z;
        g.coords(i, x, y, z);
        const double c = u[i];
        const double d0 = (double)u[g.at(x + 1 == g.nx ? 0 : x + 1, y, z)] - c;
        const double d1 = (double)u[g.at(x, y + 1 == g.ny ? 0 : y + 1, z)] - c;
        const double d2 = (double)u[g.at(x, y, z + 1 == g.nz ? 0 : z + 1)] - c;
        s += sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// per-block maxima of a volume
__global__ void k_max(const float* __restrict__ v, long long n, float* __restrict__ part) {
    __shared__ float red[RED_THREADS];
    float m = -INFINITY;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        m = fmaxf(m, v[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// v = v / s (mode 0) or v = s * v (mode 1): the normalisation around TV (Run_MSET_recon.m:267-269)
__global__ void k_scale(float* __restrict__ v, long long n, float s, int mode) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        v[i] = mode == 0 ? v[i] / s : s * v[i];
}

constexpr int RED_BLOCKS = 512;

dim3 grid_of(long long n) {
    const long long b = (n + RED_THREADS - 1) / RED_THREADS;
    return dim3((unsigned)std::max(1LL, std::min(b, 8192LL)));
}

}  // namespace

int tv_ensure(TvWork& w, long long nv) {
    if (w.nv >= nv && w.P) return 0;
    if (w.P) (void)hipFree(w.P);
    if (w.part) (void)hipFree(w.part);
    w.P = nullptr;
    w.part = nullptr;
    w.nv = 0;
    if (hipMalloc((void**)&w.P, sizeof(float) * 4 * (size_t)nv) != hipSuccess) return -1;
    if (hipMalloc((void**)&w.part, sizeof(double) * RED_BLOCKS) != hipSuccess) return -1;
    w.nv = nv;
    return 0;
}

void tv_release(TvWork& w) {
    if (w.P) (void)hipFree(w.P);
    if (w.part) (void)hipFree(w.part);
    w.P = nullptr;
    w.part = nullptr;
    w.nv = 0;
}

hipError_t tv_run(TvWork& w, const float* f, float* u, int nx, int ny, int nz, float lambda, int n_iter, double* err,
                  hipStream_t st) {
    const Grid3 g{nx, ny, nz};
    const long long nv = (long long)nx * ny * nz;
    float *P0 = w.P, *P1 = w.P + nv, *P2 = w.P + 2 * nv, *D = w.P + 3 * nv;
    const dim3 gr = grid_of(nv), bl(RED_THREADS);
    hipError_t e = launch_fill(w.P, sizeof(float) * 3 * (size_t)nv, st);   // P = 0
    if (e != hipSuccess) return e;
    for (int it = 0; it < n_iter; ++it) {
        hipLaunchKernelGGL(k_tv_dual_residual, gr, bl, 0, st, g, P0, P1, P2, f, lambda, D);
        hipLaunchKernelGGL(k_tv_project, gr, bl, 0, st, g, P0, P1, P2, D);
    }
    hipLaunchKernelGGL(k_tv_output, gr, bl, 0, st, g, P0, P1, P2, f, lambda, u);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (err) {
        hipLaunchKernelGGL(k_tv_error, dim3(RED_BLOCKS), bl, 0, st, g, u, w.part);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        std::vector<double> part(RED_BLOCKS);
        if ((e = hipMemcpyAsync(part.data(), w.part, sizeof(double) * RED_BLOCKS, hipMemcpyDeviceToHost, st)) != hipSuccess)
            return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        double s = 0.0;
        for (double p : part) s += p;
        *err = (double)lambda * s;
    }
    return hipSuccess;
}

hipError_t volume_max(TvWork& w, const float* v, long long n, float* out, hipStream_t st) {
    float* part = reinterpret_cast<float*>(w.part);      // RED_BLOCKS doubles hold RED_BLOCKS floats
    hipLaunchKernelGGL(k_max, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, v, n, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    std::vector<float> h(RED_BLOCKS);
    if ((e = hipMemcpyAsync(h.data(), part, sizeof(float) * RED_BLOCKS, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    float m = -INFINITY;
    for (float x : h) m = std::max(m, x);
    *out = m;
    return hipSuccess;
}

hipError_t volume_scale(float* v, long long n, float s, int mode, hipStream_t st) {
    hipLaunchKernelGGL(k_scale, grid_of(n), dim3(RED_THREADS), 0, st, v, n, s, mode);
    return hipGetLastError();
}

}  // namespace mset
