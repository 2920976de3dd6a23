// calibrate.hip -- known-size memory traffic for calibrating rocprofv3's FETCH_SIZE / WRITE_SIZE counters.
//
//   rocprofv3 --pmc FETCH_SIZE -- ./mset_amd/mset_calibrate [MiB]     (then the same with WRITE_SIZE)
//
// Each kernel below moves a known number of bytes once, over a buffer far larger than L2 and the MALL, so
// the counter a profiler reports for its dispatch can be divided into the known bytes.  Four access
// patterns, so that one factor is not trusted for all of them:
//   k_cal_copy   float4 read + float4 write, fully coalesced           reads S, writes S
//   k_cal_read   float4 read, one float per workgroup written           reads S, writes ~0
//   k_cal_write  float4 write only                                      reads 0, writes S
//   k_cal_tile   float2 reads and writes in 256-byte row segments of a 256 x 256 complex plane, walked
//                32 columns at a time (the column-pass pattern of the 256 x 256 plane engine)
//                                                                       reads S, writes S
// The program prints the bytes each kernel moves; it is a measurement tool, not part of the library.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                                   \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) {                                                                 \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                        \
            std::exit(1);                                                                       \
        }                                                                                       \
    } while (0)

__global__ void k_cal_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}

__global__ void k_cal_read(const float4* __restrict__ a, float* __restrict__ out, size_t n) {
    float s = 0.f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float4 v = a[i];
        s += v.x + v.y + v.z + v.w;
    }
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ void k_cal_write(float4* __restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = make_float4((float)i, 1.f, 2.f, 3.f);
}

// planes of 256 x 256 float2; workgroup g takes plane g / 8 and column tile g % 8 (32 columns), its 256
// threads walking the 256 rows 8 at a time: each row segment is 32 float2 = 256 contiguous bytes
__global__ void k_cal_tile(const float2* __restrict__ a, float2* __restrict__ b, int nplanes) {
    const int g = blockIdx.x;
    if (g >= nplanes * 8) return;
    const size_t base = (size_t)(g / 8) * 65536 + (size_t)(g % 8) * 32;
    const int c = threadIdx.x % 32, r0 = threadIdx.x / 32;
    for (int r = r0; r < 256; r += 8) {
        const size_t i = base + (size_t)r * 256 + c;
        const float2 v = a[i];
        b[i] = make_float2(v.y, v.x);
    }
}

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? (size_t)std::atoll(argv[1]) : 1024;
    const size_t bytes = mib << 20;
    const size_t n4 = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    float* out = nullptr;
    CK(hipMalloc((void**)&a, bytes));
    CK(hipMalloc((void**)&b, bytes));
    const int grid = 4096;
    CK(hipMalloc((void**)&out, grid * sizeof(float)));
    // initialise both buffers through the write kernel's sibling so the pages exist before timing
    hipLaunchKernelGGL(k_cal_write, dim3(grid), dim3(256), 0, 0, a, n4);
    CK(hipMemsetAsync(b, 0, bytes, 0));
    CK(hipDeviceSynchronize());
    hipLaunchKernelGGL(k_cal_copy, dim3(grid), dim3(256), 0, 0, a, b, n4);
    CK(hipDeviceSynchronize());
    hipLaunchKernelGGL(k_cal_read, dim3(grid), dim3(256), 0, 0, b, out, n4);
    CK(hipDeviceSynchronize());
    hipLaunchKernelGGL(k_cal_write, dim3(grid), dim3(256), 0, 0, b, n4);
    CK(hipDeviceSynchronize());
    const int nplanes = (int)(bytes / (65536 * sizeof(float2)));
    hipLaunchKernelGGL(k_cal_tile, dim3(nplanes * 8), dim3(256), 0, 0, (const float2*)a, (float2*)b, nplanes);
    CK(hipDeviceSynchronize());
    CK(hipGetLastError());
    std::printf("k_cal_copy  reads %zu writes %zu bytes\n", bytes, bytes);
    std::printf("k_cal_read  reads %zu writes %zu bytes\n", bytes, (size_t)grid * sizeof(float));
    std::printf("k_cal_write reads 0 writes %zu bytes\n", bytes);
    std::printf("k_cal_tile  reads %zu writes %zu bytes\n", (size_t)nplanes * 65536 * 8, (size_t)nplanes * 65536 * 8);
    CK(hipFree(a));
    CK(hipFree(b));
    CK(hipFree(out));
    return 0;
}
