// kernels.hpp -- declarations shared by the kernels and the host engine.
#pragma once
#include <hip/hip_runtime.h>

namespace mset {

enum { GRAD_REF = 0, GRAD_EXACT = 1 };
enum { WANT_PROBE = 1, WANT_POS = 2 };

// Per-tilt inputs and outputs of a launch that runs the positions of several tilts (a table in device
// memory; position k of tilt j is the work item (j << 24) | k).
struct TiltArgs {
    const float2* trans;         // transmission planes of the tilt's rotated volume (general routine)
    const float* tph;            // or their phases in the forward layouts (separable routine)
    const float2* probeF;        // fftn(the tilt's probe)
    const float* sqrtM;          // the tilt's patterns (sqrt, ifftshifted, transposed)
    float* xlist;                // the tilt's scan positions
    float* ylist;
    double* err_pos;             // per-position errors of the tilt
};
constexpr int ITEM_TILT_SHIFT = 24;                  // work item = (tilt << 24) | position
constexpr int ITEM_POS_MASK = (1 << ITEM_TILT_SHIFT) - 1;

// Arguments of the per-position kernels (passed by value).  Plane data is [y][x], N x N.
struct PosArgs {
    int nz, L, bin;              // volume depth, number of planes ceil(nz/bin), slice binning
    float sigma, dx;             // interaction constant, voxel size: phase = sigma * (dx * sum V)
    float cent;                  // round((N+1)/2), the 1-based centre of the scan coordinates
    float s1, s2, s3;            // step sizes (object, probe, position)
    int grad_mode;               // GRAD_REF or GRAD_EXACT
    int sset;                    // single-slice operator: L = 1, bin = nz, no propagation, phase (sigma sum V) dx
    int waves;                   // K_BATCH: WAVES_STORE, WAVES_RECOMPUTE or WAVES_CHECKPOINT (exit-wave variant)
    int ckseg;                   // WAVES_CHECKPOINT: segment length K (checkpoint_segment)
    int want;                    // WANT_PROBE | WANT_POS
    const float2* tw;            // W_N^k, k < N
    const float2* prop;          // propagator, unshifted
    const float* sqrtM;          // sqrt of the measured patterns, unshifted, [k][y][x]
    float* V;                    // K_PLAIN: volume [z][y][x], updated in place
    float2* trans;               // transmission planes [a][y][x] (K_PLAIN computes them itself)
    float2* probe;               // K_PLAIN: probe, updated in place
    const float2* probeF;        // K_BATCH / K_FWD: fftn(probe)
    float* xlist;                // scan positions (MATLAB 1-based pixels); K_PLAIN updates them
    float* ylist;
    const int* order;            // optional list of position indices (work items when tilts is set)
    const TiltArgs* tilts;       // K_FWD over several tilts: per-tilt pointers, indexed by the work item
    int first, count;            // K_BATCH / K_FWD: positions order[first .. first+count)
    int obase;                   // err_pos and dpos are indexed k - obase (0: by position index)
    double* err_pos;             // per position: sum over pixels of (sqrt M - sqrt I)^2
    float* inten;                // K_FWD, optional: |Psi|^2 per position, DC-centred (MATLAB layout)
    float2* save;                // per-slot scratch: L + 2 planes, RECOMPUTE_PLANES or checkpoint_planes(L)
    float* gslab;                // object-gradient slabs, L planes each (K_PLAIN: slab 0, this position)
    float2* pslab;               // K_BATCH: probe-gradient slabs, one plane each (Fourier domain)
    int nslab;                   // K_BATCH: position p of the launch adds to slab p % nslab
    int sep;                     // N = 256: separable-propagator routine (p1d, tph in the forward layouts,
                                 // tph_b in the adjoint layouts; odd gradient planes transposed)
    const float2* p1d;           // sep / pfac: the propagator's 1D factors px[N], py[N] (p(kx, ky) = px(kx) py(ky))
    int pfac;                    // N < 256, batched gradient with stored waves: apply the propagator as p1d's factors
    const float* tph;            // sep: transmission phases / pi, reduced to [-1, 1] (t = exp(i pi tph)), forward layouts
    const float* tph_b;          // sep, K_BATCH: the same phases in the adjoint passes' layouts
    float2* planes;              // N = 256: per-slot plane buffers (2 N^2), the tiled engine's state
    float* dpos;                 // K_BATCH: per-position update terms (2 per position)
};

// Exit-wave variants of the batched kernel.  STORE keeps every wave psi_0 .. psi_L; RECOMPUTE keeps five
// planes and rebuilds psi_a in the adjoint by inverting the forward step (unit-modulus propagator);
// CHECKPOINT keeps every K-th wave and the last segment, and rebuilds each earlier segment of K waves from
// its checkpoint by running the forward steps again (any propagator).
enum { WAVES_STORE = 0, WAVES_RECOMPUTE = 1, WAVES_CHECKPOINT = 2 };
// per-slot scratch planes of the recompute variant: psi0, Fourier psi0, two wave planes, parked adjoint
constexpr int RECOMPUTE_PLANES = 5;
// checkpoint variant: segment length K = ceil(sqrt(L + 1)) and its per-slot planes.  General routine: psi0,
// Fourier psi0, the checkpoints psi_K, psi_2K, .. psi_(L / K) K, the K - 1 other waves of one segment, the
// parked adjoint wave (the last plane).  Separable routine: psi0, the checkpoint pass inputs of the segments
// below the last ((L - 1) / K), the K waves of one segment, two planes for the rebuild passes.
__host__ __device__ inline int checkpoint_segment(int L) {
    int k = 2;
    while (k * k < L + 1) ++k;
    return k;
}
__host__ __device__ inline int checkpoint_planes(int L, int K) { return L / K + K + 3; }

bool plane_size_supported(int n);
int position_threads(int n);
hipError_t launch_plain(int n, const PosArgs& A, int k, hipStream_t st);
// shape: the workgroup shape of the batched kernel (0, or 1 = the alternative shape, stored waves only)
// shape: 0 default, 1 alternative (MSET_ENGINE_ALT_SHAPE, stored waves), 2 the separable 256 routine's
// 1024-thread shape for launches of at most one position per CU (stored or checkpointed waves)
hipError_t launch_batch(int n, const PosArgs& A, int grid, int shape, hipStream_t st);
hipError_t launch_fwd(int n, const PosArgs& A, int grid, hipStream_t st);
// forward + loss over the work items A.order[0 .. A.count) of several tilts (A.tilts)
hipError_t launch_fwd_tilts(int n, const PosArgs& A, int grid, hipStream_t st);
// 2D transform of one plane: dir 0 forward, in [y][x] -> out [x][y]; dir 1 inverse, in [x][y] -> out [y][x]
// = alpha/N^2 v + beta out, and with out2 also out2 = alpha2/N^2 v + beta2 out2.  planes: at N = 256 one
// plane of scratch (the transform runs as two passes of line transforms over the device).
hipError_t launch_plane_fft(int n, const float2* in, float2* out, const float2* tw, int dir, float alpha, float beta,
                            float2* planes, hipStream_t st, float2* out2 = nullptr, float alpha2 = 0.f,
                            float beta2 = 0.f);
bool plane_in_global(int n);     // the position kernels keep the plane in global memory (planes buffer)
hipError_t occupancy_batch(int n, int shape, int* blocks_per_cu);

// element-wise kernels (elementwise.hip)
hipError_t launch_ingest_patterns(const float* M, float* sqrtM, int n, long long npos, int transpose, hipStream_t st);
hipError_t launch_transmission(const float* V, float2* trans, int n, int nz, int bin, int sset, float sigma, float dx,
                               hipStream_t st);
hipError_t launch_fill(void* p, size_t bytes, hipStream_t st);
// separable-propagator kernels: transmission phases in the forward (Tf) and adjoint (Tb, nullable) layouts,
// and the in-place fix of the odd gradient planes they leave transposed
hipError_t launch_transmission_sep(const float* V, float* Tf, float* Tb, int n, int nz, int bin, float sigma, float dx,
                                   hipStream_t st);
hipError_t launch_untranspose_odd(float* G, int n, int L, hipStream_t st);
hipError_t launch_reduce_slabs(const float* slab, int nslab, size_t n, float* out, hipStream_t st);
hipError_t launch_reduce_cslabs(const float2* slab, int nslab, size_t n, float2* out, hipStream_t st);
hipError_t launch_update_volume(float* V, const float* G, int n, int nz, int bin, float s1, hipStream_t st);
hipError_t launch_update_positions(float* x, float* y, const float* dpos, int npos, float s3, hipStream_t st);
hipError_t launch_positivity(float* V, size_t n, hipStream_t st);

// total-variation regulariser (tv.hip): dual field (3 volumes) + scratch volume, reduction partials
struct TvWork {
    float* P = nullptr;
    double* part = nullptr;
    long long nv = 0;
};
int tv_ensure(TvWork& w, long long nv);
void tv_release(TvWork& w);
// u = TV_method(f, lambda, n_iter) on device volumes [nz][ny][nx] (u may equal f); err (nullable) gets
// lambda * TV(u)
hipError_t tv_run(TvWork& w, const float* f, float* u, int nx, int ny, int nz, float lambda, int n_iter, double* err,
                  hipStream_t st);
hipError_t volume_max(TvWork& w, const float* v, long long n, float* out, hipStream_t st);
hipError_t volume_scale(float* v, long long n, float s, int mode, hipStream_t st);

}  // namespace mset
