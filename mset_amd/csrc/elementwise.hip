// elementwise.hip -- bandwidth-bound helper kernels: pattern ingest, transmission planes, slab
// reductions, updates and positivity.  All are grid-stride loops with 16-byte-or-wider coalesced
// accesses where the layout allows; launches are sized to the element count (>= 1024 workgroups when
// the data is large) so the 256 CUs / 8 XCDs are filled.
#include <hip/hip_runtime.h>
#include "kernels.hpp"

namespace mset {

static inline dim3 grid_for(size_t n, int per_thread = 1) {
    size_t g = (n / per_thread + 255) / 256;
    if (g < 1) g = 1;
    if (g > 8192) g = 8192;
    return dim3((unsigned)g);
}

// sqrt(ifftshift(M)) of every pattern (src/CUDA/mset.cu:206-210 does this per position per visit;
// here it is done once at ingest), stored transposed for the row pass that reads it:
// out[k][x][y] = sqrt(M[k][(y + N/2) % N][(x + N/2) % N]).  transpose: each pattern is first transposed
// (STEM_data.diffraction_transpose, STEP00_INIT.m:58-73), i.e. M[k][x][y] is read instead.
__global__ void k_ingest(const float* __restrict__ M, float* __restrict__ out, int n, long long npos, int transpose) {
    const long long n2 = (long long)n * n, tot = n2 * npos;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
        const long long k = i / n2;
        const int r = (int)(i - k * n2), x = r / n, y = r % n;
        const int ys = (y + n / 2) % n, xs = (x + n / 2) % n;
        out[i] = sqrtf(M[k * n2 + (transpose ? (long long)xs * n + ys : (long long)ys * n + xs)]);
    }
}

// t_a = exp(1i * sigma * (sum_{z in bin a} V) * dx)  (STEP02_FORWARD_4D.m:12-20)
__global__ void k_trans(const float* __restrict__ V, float2* __restrict__ T, int n2, int nz, int bin, int L, int sset,
                        float sigma, float dx) {
    const long long tot = (long long)n2 * L;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
        const int a = (int)(i / n2), p = (int)(i - (long long)a * n2);
        const int z0 = a * bin, z1 = min(z0 + bin, nz);
        float s = 0.f;
        for (int z = z0; z < z1; ++z) s += V[(size_t)z * n2 + p];
        // STEP02_FORWARD_4D.m:18-19 rounds sigma * (sum * dx); STEP02_FORWARD_4D_SSET.m:9 (sigma * sum) * dx
        const float ph = sset ? (sigma * s) * dx : sigma * (s * dx);
        float sn, cs;
        sincosf(ph, &sn, &cs);
        T[i] = make_float2(cs, sn);
    }
}

// Transmission phases for the separable-propagator kernels, which form t = exp(i pi r) where they apply it
// (half the bytes of a complex plane).  r is the f32 phase of k_trans, divided by pi and reduced to [-1, 1]
// in double: t differs from k_trans's sincosf by about 1e-7.  Plane a goes to Tf in the
// layout of the pass that applies it in the forward sweep (rows, [x][y], for even a; columns, [y][x], for
// odd a) and, if Tb is given, to Tb in the layout of its adjoint pass (the other one).
__global__ void k_trans_sep(const float* __restrict__ V, float* __restrict__ Tf, float* __restrict__ Tb, int n, int nz,
                            int bin, int L, float sigma, float dx) {
    const long long n2 = (long long)n * n, tot = n2 * L;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
        const int a = (int)(i / n2), p = (int)(i - (long long)a * n2);
        const int y = p / n, x = p % n;
        const int z0 = a * bin, z1 = min(z0 + bin, nz);
        float s = 0.f;
        for (int z = z0; z < z1; ++z) s += V[(size_t)z * n2 + p];
        const float ph = sigma * (s * dx);                      // STEP02_FORWARD_4D.m:18-19 rounding order
        double r = (double)ph / M_PI;
        r -= 2.0 * rint(0.5 * r);                               // [-1, 1]
        const long long rowmaj = (long long)y * n + x, colmaj = (long long)x * n + y;
        Tf[(long long)a * n2 + ((a & 1) ? rowmaj : colmaj)] = (float)r;
        if (Tb) Tb[(long long)a * n2 + ((a & 1) ? colmaj : rowmaj)] = (float)r;
    }
}

// The separable kernels leave odd gradient planes transposed ([x][y]); swap them back to [y][x] in place.
__global__ void k_untranspose_odd(float* __restrict__ G, int n, int L) {
    const long long n2 = (long long)n * n, tot = n2 * (L / 2);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
        const int a = 2 * (int)(i / n2) + 1;
        const int p = (int)(i % n2), y = p / n, x = p % n;
        if (x < y) {
            float* P = G + (size_t)a * n2;
            const float u = P[(size_t)y * n + x], w = P[(size_t)x * n + y];
            P[(size_t)y * n + x] = w;
            P[(size_t)x * n + y] = u;
        }
    }
}

__global__ void k_fill(uint4* p, size_t n16) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0, 0, 0, 0);
}
__global__ void k_fill_bytes(unsigned char* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

// out[i] = sum_s slab[s][i], s ascending (fixed order: bitwise repeatable)
// Slabs are summed in slot order (s = 0, 1, ...), so the result is the same for every launch shape;
// U loads are in flight at a time.
__global__ void k_reduce(const float4* __restrict__ slab, int nslab, size_t n4, float4* __restrict__ out) {
    constexpr int U = 8;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
        int s = 0;
        for (; s + U <= nslab; s += U) {
            float4 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = slab[(size_t)(s + u) * n4 + i];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                a.x += v[u].x; a.y += v[u].y; a.z += v[u].z; a.w += v[u].w;
            }
        }
        for (; s < nslab; ++s) {
            const float4 v = slab[(size_t)s * n4 + i];
            a.x += v.x; a.y += v.y; a.z += v.z; a.w += v.w;
        }
        out[i] = a;
    }
}
__global__ void k_reduce1(const float* __restrict__ slab, int nslab, size_t n, float* __restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float a = 0.f;
        for (int s = 0; s < nslab; ++s) a += slab[(size_t)s * n + i];
        out[i] = a;
    }
}

// V[z] += s1 * G[bin(z)]  (the per-bin gradient replicated over the bin, STEP03_BACK.m:17-21)
__global__ void k_update_volume(float* __restrict__ V, const float* __restrict__ G, int n2, int nz, int bin, float s1) {
    const long long tot = (long long)n2 * nz;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i / n2), p = (int)(i - (long long)z * n2);
        V[i] += s1 * G[(size_t)(z / bin) * n2 + p];
    }
}

__global__ void k_update_positions(float* x, float* y, const float* __restrict__ d, int npos, float s3) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < npos; k += gridDim.x * blockDim.x) {
        x[k] = x[k] + s3 * d[2 * k];
        y[k] = y[k] + s3 * d[2 * k + 1];
    }
}

// rec(rec < 0) = 0  (Run_MSET_recon.m:183-185)
__global__ void k_positivity(float* V, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        V[i] = V[i] < 0.f ? 0.f : V[i];
}

hipError_t launch_ingest_patterns(const float* M, float* sqrtM, int n, long long npos, int transpose, hipStream_t st) {
    hipLaunchKernelGGL(k_ingest, grid_for((size_t)n * n * npos), dim3(256), 0, st, M, sqrtM, n, npos, transpose);
    return hipGetLastError();
}

hipError_t launch_transmission(const float* V, float2* T, int n, int nz, int bin, int sset, float sigma, float dx,
                               hipStream_t st) {
    const int L = (nz + bin - 1) / bin;
    hipLaunchKernelGGL(k_trans, grid_for((size_t)n * n * L), dim3(256), 0, st, V, T, n * n, nz, bin, L, sset, sigma, dx);
    return hipGetLastError();
}

hipError_t launch_transmission_sep(const float* V, float* Tf, float* Tb, int n, int nz, int bin, float sigma, float dx,
                                   hipStream_t st) {
    const int L = (nz + bin - 1) / bin;
    hipLaunchKernelGGL(k_trans_sep, grid_for((size_t)n * n * L), dim3(256), 0, st, V, Tf, Tb, n, nz, bin, L, sigma, dx);
    return hipGetLastError();
}

hipError_t launch_untranspose_odd(float* G, int n, int L, hipStream_t st) {
    if (L < 2) return hipSuccess;
    hipLaunchKernelGGL(k_untranspose_odd, grid_for((size_t)n * n * (L / 2)), dim3(256), 0, st, G, n, L);
    return hipGetLastError();
}

hipError_t launch_fill(void* p, size_t bytes, hipStream_t st) {
    if (((uintptr_t)p % 16) == 0 && bytes % 16 == 0)
        hipLaunchKernelGGL(k_fill, grid_for(bytes / 16), dim3(256), 0, st, (uint4*)p, bytes / 16);
    else
        hipLaunchKernelGGL(k_fill_bytes, grid_for(bytes), dim3(256), 0, st, (unsigned char*)p, bytes);
    return hipGetLastError();
}

hipError_t launch_reduce_slabs(const float* slab, int nslab, size_t n, float* out, hipStream_t st) {
    if (n % 4 == 0 && ((uintptr_t)slab % 16) == 0 && ((uintptr_t)out % 16) == 0)
        hipLaunchKernelGGL(k_reduce, grid_for(n / 4), dim3(256), 0, st, (const float4*)slab, nslab, n / 4, (float4*)out);
    else
        hipLaunchKernelGGL(k_reduce1, grid_for(n), dim3(256), 0, st, slab, nslab, n, out);
    return hipGetLastError();
}

hipError_t launch_reduce_cslabs(const float2* slab, int nslab, size_t n, float2* out, hipStream_t st) {
    return launch_reduce_slabs((const float*)slab, nslab, 2 * n, (float*)out, st);
}

hipError_t launch_update_volume(float* V, const float* G, int n, int nz, int bin, float s1, hipStream_t st) {
    hipLaunchKernelGGL(k_update_volume, grid_for((size_t)n * n * nz), dim3(256), 0, st, V, G, n * n, nz, bin, s1);
    return hipGetLastError();
}

hipError_t launch_update_positions(float* x, float* y, const float* dpos, int npos, float s3, hipStream_t st) {
    hipLaunchKernelGGL(k_update_positions, grid_for((size_t)npos), dim3(256), 0, st, x, y, dpos, npos, s3);
    return hipGetLastError();
}

hipError_t launch_positivity(float* V, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_positivity, grid_for(n), dim3(256), 0, st, V, n);
    return hipGetLastError();
}

}  // namespace mset
