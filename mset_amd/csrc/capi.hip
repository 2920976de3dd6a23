// capi.hip -- host engine behind include/mset.h.
//
// mset_ctx holds the device state of one tilt (the role of mset::paras, src/CUDA/mset.cuh:23-115, and
// of the module-static shared_ptr of src/CUDA/mex_main.cu:13).  The large per-workgroup scratch of
// the batched kernels lives in a Workspace that the tilts of one reconstruction share.  mset_recon
// keeps a whole tilt series resident and runs the update and error passes of
// src/Run_MSET_recon.m:119-239 on the device, optionally replaying a captured hipGraph.
// Every HIP call is checked; failures become status codes plus a thread-local message instead of
// mexErrMsgTxt or unchecked calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mset.h"
#include "kernels.hpp"
#include "rotate.hpp"

using namespace mset;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(e_ == hipErrorOutOfMemory ? MSET_ENOMEM : MSET_EDEVICE,                \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                    \
    } while (0)

#define CHK(expr)                         \
    do {                                  \
        int r_ = (expr);                  \
        if (r_ != MSET_OK) return r_;     \
    } while (0)

template <typename T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int ensure(size_t count, const char* what) {
        if (count <= n && p) return MSET_OK;
        release();
        if (count == 0) return MSET_OK;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(MSET_ENOMEM, std::string("hipMalloc ") + what + ": " + hipGetErrorString(e));
        }
        n = count;
        return MSET_OK;
    }
};

struct Prof {
    int launches = 0;
    double ms = 0.0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};
using ProfMap = std::map<std::string, Prof>;

// scratch shared by the tilts of one reconstruction
struct Workspace {
    DevBuf<float2> save, trans, probeF, pslab, PG, planes;
    DevBuf<float> gslab, G;
    int exit_waves = MSET_EXIT_WAVES_AUTO;   // requested exit-wave variant of the batched kernels
    int slot_limit = 0;                      // cap on the persistent grid of the batched kernels (0: none)
    int engine = MSET_ENGINE_AUTO;           // MSET_ENGINE_* flags: plane routine at 256, workgroup shape
    int shape = 0;                           // workgroup shape of the last batched call (Cfg<N, shape>)
    int waves = WAVES_STORE;                 // the variant of the last batched call (WAVES_*)
    size_t slot_bytes = 0;                   // its per-slot scratch (waves + gradient slab)
    RotWork rot;
    TvWork tv;
    ~Workspace() {
        rot_release(rot);
        tv_release(tv);
    }
};

}  // namespace

struct mset_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;
    int cus = 0;
    mset_params p{};
    int n = 0, nz = 0, L = 0, npos = 0;
    int method = MSET_METHOD_MSET;   // operator of the current call (MSET, or SSET = one plane of depth nz)
    bool uploaded = false;
    bool prop_unitary = false;       // |prop| = 1 everywhere (to 1e-5): the forward sweep can be inverted
    bool prop_sep = false;           // prop(kx, ky) = px(kx) py(ky) (to 2e-6): the separable kernels apply
    DevBuf<float2> p1d;              // px, py (N each) when prop_sep
    DevBuf<float> V, sqrtM, x, y, dpos, Mtmp;
    DevBuf<float2> probe, prop, tw;
    DevBuf<double> err;
    DevBuf<float2> mt_trans, mt_probeF;   // this tilt's transmission planes and fftn(probe) in a multi-tilt launch
    std::shared_ptr<Workspace> ws;
    int slots = 0;
    bool* profiling = nullptr;
    std::shared_ptr<ProfMap> prof;
    bool own_profiling = false;
};

namespace {

// event bracket around a launch when profiling is on (never inside a graph capture)
struct Timed {
    hipStream_t st;
    ProfMap* pm;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(mset_ctx* c, const char* n_) : st(c->stream), pm(*c->profiling ? c->prof.get() : nullptr), name(n_) {
        if (pm) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, st);
        }
    }
    ~Timed() {
        if (pm) {
            (void)hipEventRecord(b, st);
            (*pm)[name].pending.emplace_back(a, b);
        }
    }
};

int prof_read(ProfMap& pm, hipStream_t st, const char* name, int* launches, double* total_ms) {
    HIPCHK(hipStreamSynchronize(st));
    Prof& P = pm[name];
    for (auto& ev : P.pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev.first, ev.second));
        P.ms += ms;
        P.launches += 1;
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    P.pending.clear();
    if (launches) *launches = P.launches;
    if (total_ms) *total_ms = P.ms;
    P.launches = 0;
    P.ms = 0.0;
    return MSET_OK;
}

void prof_free(ProfMap& pm) {
    for (auto& kv : pm)
        for (auto& ev : kv.second.pending) {
            (void)hipEventDestroy(ev.first);
            (void)hipEventDestroy(ev.second);
        }
    pm.clear();
}

// Scratch of mset_grad_noalloc: a file-scope device array, so that entry point never allocates.  Each GPU
// holds its own copy at its own address; host code reaches it only through the address that
// hipGetSymbolAddress returns on that device (the array's name is a host-side stand-in).
__device__ __align__(256) unsigned char scratch[MSET_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch;   // device -> address of `scratch` there

// resolves the scratch address on `device` once; the device must be current
int resolve_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch.count(device)) return MSET_OK;
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(MSET_EDEVICE, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch[device] = (unsigned char*)ptr;
    return MSET_OK;
}

unsigned char* scratch_on(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(device);
    return it == g_scratch.end() ? nullptr : it->second;
}

int check_ctx(mset_ctx* c) {
    if (!c) return fail(MSET_EINVAL, "null context");
    HIPCHK(hipSetDevice(c->device));
    return MSET_OK;
}

int check_uploaded(mset_ctx* c) {
    CHK(check_ctx(c));
    if (!c->uploaded) return fail(MSET_EINVAL, "no data uploaded (call mset_upload first)");
    return MSET_OK;
}

// selects the operator of the following work on the context: 'run_mset' / 'run_sset' (mex_main.cu:26-35)
int set_method(mset_ctx* c, int method) {
    if (method != MSET_METHOD_MSET && method != MSET_METHOD_SSET)
        return fail(MSET_EINVAL, "method must be MSET_METHOD_MSET or MSET_METHOD_SSET");
    c->method = method;
    return MSET_OK;
}

void twiddles(int n, std::vector<float2>& tw) {
    tw.resize(n);
    for (int k = 0; k < n; ++k) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        tw[k] = make_float2((float)cosl(a), (float)sinl(a));
    }
}

size_t plane(const mset_ctx* c) { return (size_t)c->n * c->n; }
bool sset(const mset_ctx* c) { return c->method == MSET_METHOD_SSET; }
// planes and slices per plane of the current operator: SSET projects the whole depth into one plane
// (STEP02_FORWARD_4D_SSET.m:9)
int planes(const mset_ctx* c) { return sset(c) ? 1 : c->L; }
int binning(const mset_ctx* c) { return sset(c) ? c->nz : c->p.slice_binning; }

PosArgs base_args(mset_ctx* c) {
    PosArgs A{};
    A.nz = c->nz;
    A.L = planes(c);
    A.bin = binning(c);
    A.sset = sset(c) ? 1 : 0;
    A.sigma = c->p.sigma;
    A.dx = c->p.potential_pixelsize;
    A.cent = (float)std::floor((c->n + 1) / 2.0 + 0.5);
    A.s1 = c->p.step_size[0];
    A.s2 = c->p.step_size[1];
    A.s3 = c->p.step_size[2];
    A.grad_mode = c->p.grad_mode;
    A.want = (c->p.step_size[1] != 0.f ? WANT_PROBE : 0) | (c->p.step_size[2] != 0.f ? WANT_POS : 0);
    A.tw = c->tw.p;
    A.prop = c->prop.p;
    A.sqrtM = c->sqrtM.p;
    A.V = c->V.p;
    A.trans = c->ws->trans.p;
    A.probe = c->probe.p;
    A.probeF = c->ws->probeF.p;
    A.xlist = c->x.p;
    A.ylist = c->y.p;
    A.err_pos = c->err.p;
    A.save = c->ws->save.p;
    A.waves = c->ws->waves;
    A.ckseg = checkpoint_segment(planes(c));
    A.gslab = c->ws->gslab.p;
    A.pslab = c->ws->pslab.p;
    A.dpos = c->dpos.p;
    A.planes = plane_in_global(c->n) ? c->ws->planes.p : nullptr;
    A.sep = 0;
    A.p1d = c->p1d.p;
    A.pfac = c->prop_sep && !(c->ws->engine & MSET_ENGINE_GENERAL) ? 1 : 0;
    // the separable routine's phases share the transmission buffer: forward layouts, then adjoint layouts
    A.tph = (const float*)c->ws->trans.p;
    A.tph_b = A.tph + (size_t)c->L * plane(c);
    return A;
}

// The separable-propagator kernels (the 256 x 256 engine) serve the forward-only launches and the batched
// gradient with the reference pairing and stored waves; everything else runs the general kernels.
bool sep_fwd(const mset_ctx* c) {
    return plane_in_global(c->n) && c->prop_sep && !sset(c) && !(c->ws->engine & MSET_ENGINE_GENERAL);
}
bool sep_batch(const mset_ctx* c) {
    return sep_fwd(c) && c->p.grad_mode == GRAD_REF && (c->ws->waves == WAVES_STORE || c->ws->waves == WAVES_CHECKPOINT);
}

// workgroup shape of a batched call: the alternative shape (MSET_ENGINE_ALT_SHAPE) exists for stored waves
int shape_for(const mset_ctx* c, int waves) {
    return (c->ws->engine & MSET_ENGINE_ALT_SHAPE) && waves == WAVES_STORE ? 1 : 0;
}

// workgroups of the batched kernel of a workgroup shape the device keeps resident at once (1 per CU at
// N = 128 in the default shape); by default the shape of the last batched call
int device_slots(mset_ctx* c, int shape) {
    int per_cu = 1;
    if (occupancy_batch(c->n, shape, &per_cu) != hipSuccess || per_cu < 1) per_cu = 1;
    return c->cus * per_cu;
}
int device_slots(mset_ctx* c) { return device_slots(c, c->ws->shape); }

// the persistent grid of the batched kernels: the resident workgroups, capped by the slot limit and the
// positions
int batch_slots(mset_ctx* c, int shape) {
    int s = device_slots(c, shape);
    if (c->ws->slot_limit > 0 && s > c->ws->slot_limit) s = c->ws->slot_limit;
    if (s > c->npos) s = c->npos;
    return s < 1 ? 1 : s;
}
int batch_slots(mset_ctx* c) { return batch_slots(c, c->ws->shape); }

// Gradient slabs of the context's batched launches: position p of a launch adds to slab p % S.  S depends
// on the device, the workgroup shape and the positions only, never on the slot limit, so capping the grid
// leaves every sum (and so every result) bitwise unchanged.
int grad_slabs(mset_ctx* c, int shape) { return std::max(1, std::min(device_slots(c, shape), c->npos)); }
int grad_slabs(mset_ctx* c) { return grad_slabs(c, c->ws->shape); }

// grid of a batched launch of `count` positions on S slabs: at most `want` workgroups and, when slabs take
// several positions (count > S), a divisor of S, so that each slab's positions run on one workgroup in order
// The workgroup shape of a batched launch: a launch of the separable 256 routine (stored or checkpointed
// waves) with at most one position per CU (a tilt's last mini-batch) runs the 1024-thread shape, one
// workgroup of 16 waves per CU instead of one or two of 8 (15-20 % shorter at 64-256 positions).  Its line
// transforms, operand order and slab assignment are the default shape's, so the gradient slabs are the
// same bit for bit; only the workgroup sums (loss, position terms) add their terms in another order.
int batch_shape(const mset_ctx* c, int shape, bool sep, int waves, int count) {
    const bool wide = waves == WAVES_STORE || waves == WAVES_CHECKPOINT;
    return c->n == 256 && shape == 0 && sep && wide && count <= c->cus ? 2 : shape;
}

int launch_grid(int want, int count, int S) {
    int g = std::max(1, std::min(want, count));
    if (count > S)
        while (S % g) --g;
    return g;
}

int ensure_fixed_buffers(mset_ctx* c) {
    Workspace& W = *c->ws;
    // complex planes, or (separable routine) f32 phases in the forward and the adjoint layouts: L planes
    // either way
    CHK(W.trans.ensure(plane(c) * c->L, "transmission"));
    CHK(W.probeF.ensure(plane(c), "probe spectrum"));
    // the position kernels of the largest planes keep the plane in global memory: 2 planes per slot
    if (plane_in_global(c->n)) CHK(W.planes.ensure((size_t)batch_slots(c) * 2 * plane(c), "plane buffers"));
    return MSET_OK;
}

// exit-wave variant of a batched call (see mset_set_exit_waves).  AUTO stores unless the store scratch would
// exceed a quarter of the device; then it checkpoints if that scratch fits (faster than recomputing and the
// stored variant's waves), else recomputes (unit-modulus propagator), else checkpoints.  RECOMPUTE applies
// from 4 planes on; the single-slice operator always stores.
int wave_planes(const mset_ctx* c, int waves) {
    const int L = planes(c);
    return waves == WAVES_RECOMPUTE ? RECOMPUTE_PLANES
           : waves == WAVES_CHECKPOINT ? checkpoint_planes(L, checkpoint_segment(L)) : L + 2;
}

int resolve_exit_waves(mset_ctx* c, int slots, int* waves) {
    Workspace& W = *c->ws;
    const int L = planes(c);
    *waves = WAVES_STORE;
    if (sset(c)) return MSET_OK;
    if (W.exit_waves == MSET_EXIT_WAVES_RECOMPUTE) {
        if (!c->prop_unitary) return fail(MSET_EINVAL, "recomputed exit waves need a unit-modulus propagator");
        if (L + 2 > RECOMPUTE_PLANES) *waves = WAVES_RECOMPUTE;
    } else if (W.exit_waves == MSET_EXIT_WAVES_CHECKPOINT) {
        *waves = WAVES_CHECKPOINT;
    } else if (W.exit_waves == MSET_EXIT_WAVES_AUTO) {
        size_t fr = 0, total = 0;
        HIPCHK(hipMemGetInfo(&fr, &total));
        auto scratch = [&](int wv) {
            return (double)slots * (double)plane(c) * (wave_planes(c, wv) * sizeof(float2) + L * sizeof(float));
        };
        const double quarter = 0.25 * (double)total;
        if (scratch(WAVES_STORE) > quarter) {
            int alt = WAVES_CHECKPOINT;
            if (scratch(WAVES_CHECKPOINT) > quarter && c->prop_unitary && L + 2 > RECOMPUTE_PLANES) alt = WAVES_RECOMPUTE;
            if (L + 2 > wave_planes(c, alt)) *waves = alt;
        }
    }
    return MSET_OK;
}

int ensure_batch_buffers(mset_ctx* c) {
    Workspace& W = *c->ws;
    int wv = WAVES_STORE;
    CHK(resolve_exit_waves(c, batch_slots(c, 0), &wv));
    W.waves = wv;
    W.shape = shape_for(c, wv);
    const int s = batch_slots(c);
    CHK(ensure_fixed_buffers(c));
    const int save_planes = wave_planes(c, wv);
    CHK(W.save.ensure((size_t)s * save_planes * plane(c), "save waves"));
    W.slot_bytes = plane(c) * (save_planes * sizeof(float2) + (size_t)planes(c) * sizeof(float));
    const int S = grad_slabs(c);
    CHK(W.gslab.ensure((size_t)S * c->L * plane(c), "gradient slabs"));
    CHK(W.pslab.ensure((size_t)S * plane(c), "probe slabs"));
    CHK(W.G.ensure((size_t)c->L * plane(c), "gradient"));
    CHK(W.PG.ensure(plane(c), "probe gradient"));
    CHK(c->dpos.ensure((size_t)2 * c->npos, "position terms"));
    c->slots = s;
    return MSET_OK;
}

// transmission planes + fftn(probe) for the fixed-state kernels; sep: 1 = forward layouts of the separable
// kernels, 2 = forward and adjoint layouts
int prepare_fixed(mset_ctx* c, int sep = 0) {
    CHK(ensure_fixed_buffers(c));
    if (sep)
        HIPCHK(launch_transmission_sep(c->V.p, (float*)c->ws->trans.p, sep == 2 ? (float*)c->ws->trans.p + plane(c) * c->L : nullptr,
                                       c->n, c->nz, binning(c), c->p.sigma, c->p.potential_pixelsize, c->stream));
    else
        HIPCHK(launch_transmission(c->V.p, c->ws->trans.p, c->n, c->nz, binning(c), sset(c) ? 1 : 0, c->p.sigma,
                                   c->p.potential_pixelsize, c->stream));
    HIPCHK(launch_plane_fft(c->n, c->probe.p, c->ws->probeF.p, c->tw.p, 0, 1.f, 0.f, c->ws->planes.p, c->stream));
    return MSET_OK;
}

double host_error_sum(const mset_ctx* c, const std::vector<double>& e) {
    double s = 0.0;
    for (double v : e) s += v;
    return s / ((double)c->p.n_scan_x * (double)c->p.n_scan_y) / (double)plane(c);
}

int read_errors(mset_ctx* c, int first, int count, double* out) {
    std::vector<double> e(count);
    HIPCHK(hipMemcpyAsync(e.data(), c->err.p + first, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *out = host_error_sum(c, e);
    return MSET_OK;
}

// The sequential path on the separable 256 routine: each position is a batched launch of that one position
// (the 1024-thread shape, one workgroup) with its gradient planes, probe term and position terms, and the
// host applies the three steps before the next position's transmission and fftn(probe).  The semantics of
// the one-position kernel below, at one plane pass per propagation step instead of two.
int issue_plain_sep(mset_ctx* c) {
    Workspace& W = *c->ws;
    CHK(W.pslab.ensure(plane(c), "probe term"));
    CHK(c->dpos.ensure((size_t)2 * c->npos, "position terms"));
    PosArgs A = base_args(c);
    A.sep = 1;
    A.waves = WAVES_STORE;
    A.count = 1;
    A.nslab = 1;
    const float* s = c->p.step_size;
    for (int k = 0; k < c->npos; ++k) {
        HIPCHK(launch_transmission_sep(c->V.p, (float*)W.trans.p, (float*)W.trans.p + plane(c) * c->L, c->n, c->nz,
                                       binning(c), c->p.sigma, c->p.potential_pixelsize, c->stream));
        HIPCHK(launch_plane_fft(c->n, c->probe.p, W.probeF.p, c->tw.p, 0, 1.f, 0.f, W.planes.p, c->stream));
        A.first = k;
        {
            Timed t(c, "position_plain");
            HIPCHK(launch_batch(c->n, A, 1, 2, c->stream));
        }
        // the gradient planes (odd planes transposed by the routine), then the steps (STEP03_BACK.m:84-92)
        HIPCHK(launch_untranspose_odd(W.gslab.p, c->n, c->L, c->stream));
        if (s[0] != 0.f) HIPCHK(launch_update_volume(c->V.p, W.gslab.p, c->n, c->nz, binning(c), s[0], c->stream));
        if (s[1] != 0.f)
            HIPCHK(launch_plane_fft(c->n, W.pslab.p, c->probe.p, c->tw.p, 1, -s[1], 1.f, W.planes.p, c->stream));
        if (s[2] != 0.f)
            HIPCHK(launch_update_positions(c->x.p + k, c->y.p + k, c->dpos.p + 2 * (size_t)k, 1, s[2], c->stream));
    }
    return MSET_OK;
}

// one launch per position, sequential updates (issue only; no host synchronisation)
int issue_plain(mset_ctx* c) {
    CHK(ensure_fixed_buffers(c));
    CHK(c->ws->save.ensure((size_t)(c->L + 2) * plane(c), "save waves"));
    CHK(c->ws->gslab.ensure((size_t)c->L * plane(c), "gradient planes"));
    if (sep_fwd(c) && c->p.grad_mode == GRAD_REF) return issue_plain_sep(c);
    PosArgs A = base_args(c);
    const float s1 = c->p.step_size[0];
    for (int k = 0; k < c->npos; ++k) {
        // transmission planes from the current volume, the position, then its object step
        // (STEP02_FORWARD_4D.m:12-20, STEP03_BACK.m:84): the volume passes run on the whole device
        HIPCHK(launch_transmission(c->V.p, c->ws->trans.p, c->n, c->nz, binning(c), sset(c) ? 1 : 0, c->p.sigma,
                                   c->p.potential_pixelsize, c->stream));
        {
            Timed t(c, "position_plain");
            HIPCHK(launch_plain(c->n, A, k, c->stream));
        }
        if (s1 != 0.f) HIPCHK(launch_update_volume(c->V.p, c->ws->gslab.p, c->n, c->nz, binning(c), s1, c->stream));
    }
    return MSET_OK;
}

int issue_fwd(mset_ctx* c, float* inten) {
    const bool sep = sep_fwd(c);
    CHK(prepare_fixed(c, sep ? 1 : 0));
    PosArgs A = base_args(c);
    A.sep = sep ? 1 : 0;
    A.first = 0;
    A.count = c->npos;
    A.inten = inten;
    Timed t(c, "position_fwd");
    HIPCHK(launch_fwd(c->n, A, batch_slots(c), c->stream));
    return MSET_OK;
}

// gradient sums at the fixed state over [first, first+count) into W.G (L planes), W.PG (probe step
// direction, real space) and c->dpos; with apply, the steps are taken (RVol += s1 G, probe -= s2 PG,
// positions += s3 dpos).  want: the terms to form (the gradient request forms all of them).
int issue_grad(mset_ctx* c, int first, int count, bool apply, int want) {
    CHK(ensure_batch_buffers(c));
    const bool sep = sep_batch(c);
    CHK(prepare_fixed(c, sep ? 2 : 0));
    Workspace& W = *c->ws;
    const int S = grad_slabs(c);
    const int grid = launch_grid(c->slots, count, S);
    const int used = std::min(count, S);                 // slabs written by this launch
    PosArgs A = base_args(c);
    A.first = first;
    A.count = count;
    A.want = want;
    A.nslab = S;
    A.sep = sep ? 1 : 0;
    // every used slab receives at least one position, and its first one writes it
    {
        Timed t(c, "position_batch");
        HIPCHK(launch_batch(c->n, A, grid, batch_shape(c, W.shape, sep, W.waves, count), c->stream));
    }
    HIPCHK(launch_reduce_slabs(W.gslab.p, used, (size_t)planes(c) * plane(c), W.G.p, c->stream));
    if (sep) HIPCHK(launch_untranspose_odd(W.G.p, c->n, c->L, c->stream));
    const bool probe_step = apply && c->p.step_size[1] != 0.f && (want & WANT_PROBE);
    if (want & WANT_PROBE) {
        // in place into slab 0 (each element is read and written by the same thread), then PG = ifftn and,
        // with a probe step, probe -= s2 PG from the same transform
        HIPCHK(launch_reduce_cslabs(W.pslab.p, used, plane(c), W.pslab.p, c->stream));
        HIPCHK(launch_plane_fft(c->n, W.pslab.p, W.PG.p, c->tw.p, 1, 1.f, 0.f, W.planes.p, c->stream,
                                probe_step ? c->probe.p : nullptr, -c->p.step_size[1], 1.f));
    }
    if (apply) {
        if (c->p.step_size[0] != 0.f)
            HIPCHK(launch_update_volume(c->V.p, W.G.p, c->n, c->nz, binning(c), c->p.step_size[0], c->stream));
        if (c->p.step_size[2] != 0.f && (want & WANT_POS))
            HIPCHK(launch_update_positions(c->x.p + first, c->y.p + first, c->dpos.p + 2 * (size_t)first, count,
                                           c->p.step_size[2], c->stream));
    }
    return MSET_OK;
}

int want_from_steps(const mset_ctx* c) {
    return (c->p.step_size[1] != 0.f ? WANT_PROBE : 0) | (c->p.step_size[2] != 0.f ? WANT_POS : 0);
}

int issue_batched(mset_ctx* c, int batch) {
    for (int first = 0; first < c->npos; first += batch) {
        const int cnt = std::min(batch, c->npos - first);
        CHK(issue_grad(c, first, cnt, true, want_from_steps(c)));
    }
    return MSET_OK;
}

int new_ctx(mset_ctx** out, int device, hipStream_t shared_stream, std::shared_ptr<Workspace> ws,
            bool* profiling, std::shared_ptr<ProfMap> prof) {
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    CHK(resolve_scratch(device));
    mset_ctx* c = new mset_ctx();
    c->device = device;
    c->cus = prop.multiProcessorCount;
    if (shared_stream) {
        c->stream = shared_stream;
        c->own_stream = false;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return fail(MSET_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
    }
    c->ws = ws ? ws : std::make_shared<Workspace>();
    if (profiling) {
        c->profiling = profiling;
    } else {
        c->profiling = new bool(false);
        c->own_profiling = true;
    }
    c->prof = prof ? prof : std::make_shared<ProfMap>();
    *out = c;
    return MSET_OK;
}

void free_ctx(mset_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->prof.use_count() == 1) prof_free(*c->prof);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    if (c->own_profiling) delete c->profiling;
    delete c;
}

// validate + set geometry + copy a tilt's inputs (MEX 'upload', mex_main.cu:42-95)
int upload_impl(mset_ctx* c, const mset_params* p, const float* RVol, const float* init_wave2D, const float* prop2D,
                const float* xlist, const float* ylist, const float* measured) {
    if (!p || !init_wave2D || !prop2D || !xlist || !ylist || !measured) return fail(MSET_EINVAL, "null input array");
    if (p->nx != p->ny) return fail(MSET_EINVAL, "RVol must be square in x, y (nx == ny)");
    if (p->nx <= 0 || p->nz <= 0 || p->num_scan_pos <= 0 || p->n_scan_x <= 0 || p->n_scan_y <= 0)
        return fail(MSET_EINVAL, "sizes must be positive");
    if (!plane_size_supported(p->nx))
        return fail(MSET_EUNSUPPORTED, "lateral size " + std::to_string(p->nx) +
                                           " not built (supported: 16, 24, 32, 48, 64, 96, 128, 256)");
    if (p->slice_binning <= 0) return fail(MSET_EINVAL, "slice_binning must be a positive integer");
    if (p->grad_mode != MSET_GRAD_REF && p->grad_mode != MSET_GRAD_EXACT) return fail(MSET_EINVAL, "bad grad_mode");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->uploaded = false;
    c->p = *p;
    if (c->p.slice_binning > p->nz) c->p.slice_binning = p->nz;   // STEP00_INIT.m:127-129
    c->n = p->nx;
    c->nz = p->nz;
    c->L = (p->nz + c->p.slice_binning - 1) / c->p.slice_binning;
    c->npos = p->num_scan_pos;
    const size_t n2 = plane(c);
    CHK(c->V.ensure(n2 * c->nz, "volume"));
    CHK(c->probe.ensure(n2, "probe"));
    CHK(c->prop.ensure(n2, "propagator"));
    CHK(c->tw.ensure(c->n, "twiddles"));
    CHK(c->sqrtM.ensure(n2 * c->npos, "patterns"));
    CHK(c->x.ensure(c->npos, "scan x"));
    CHK(c->y.ensure(c->npos, "scan y"));
    CHK(c->err.ensure(c->npos, "errors"));
    CHK(c->Mtmp.ensure(n2 * c->npos, "pattern staging"));

    // ifftshift of the MATLAB-layout propagator (mex_main.cu:80), stored transposed ([x][y]) because
    // the kernels apply it in the row pass that ends each forward transform
    std::vector<float2> pu(n2);
    const int n = c->n, h = n / 2;
    for (int y = 0; y < n; ++y)
        for (int xx = 0; xx < n; ++xx) {
            const size_t src = (size_t)((y + h) % n) * n + (xx + h) % n;
            pu[(size_t)xx * n + y] = make_float2(prop2D[2 * src], prop2D[2 * src + 1]);
        }
    double dev = 0.0;
    for (const float2& q : pu) dev = std::max(dev, std::fabs((double)q.x * q.x + (double)q.y * q.y - 1.0));
    c->prop_unitary = dev <= 1e-5;
    // Separable propagator (Func_generate_propagator.m: exp(-i pi lambda dz (qx^2 + qy^2))): with p(ky, kx) =
    // pu[kx][ky], px(kx) = p(0, kx) and py(ky) = p(ky, 0) / p(0, 0) reproduce every element to FP32 rounding
    // (f1 stays alive until the synchronisation below: its copy is asynchronous)
    std::vector<float2> f1(2 * (size_t)n);
    {
        const double r00 = pu[0].x, i00 = pu[0].y, m00 = r00 * r00 + i00 * i00;
        bool sep = m00 > 0.25;
        for (int kx = 0; kx < n; ++kx) f1[kx] = pu[(size_t)kx * n];
        for (int ky = 0; ky < n && sep; ++ky) {
            const double a = pu[ky].x, b = pu[ky].y;                      // p(ky, 0) / p(0, 0)
            f1[n + ky] = make_float2((float)((a * r00 + b * i00) / m00), (float)((b * r00 - a * i00) / m00));
        }
        double worst = 0.0;
        for (int kx = 0; kx < n && sep; ++kx)
            for (int ky = 0; ky < n; ++ky) {
                const float2 u = f1[kx], w = f1[n + ky], q = pu[(size_t)kx * n + ky];
                const double re = (double)u.x * w.x - (double)u.y * w.y, im = (double)u.x * w.y + (double)u.y * w.x;
                worst = std::max(worst, std::hypot(re - q.x, im - q.y));
            }
        c->prop_sep = sep && worst <= 2e-6;
        if (c->prop_sep) {
            CHK(c->p1d.ensure(2 * (size_t)n, "propagator factors"));
            HIPCHK(hipMemcpyAsync(c->p1d.p, f1.data(), sizeof(float2) * 2 * n, hipMemcpyHostToDevice, c->stream));
        }
    }
    std::vector<float2> tw;
    twiddles(n, tw);
    if (RVol) HIPCHK(hipMemcpyAsync(c->V.p, RVol, sizeof(float) * n2 * c->nz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->probe.p, init_wave2D, sizeof(float2) * n2, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->prop.p, pu.data(), sizeof(float2) * n2, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->tw.p, tw.data(), sizeof(float2) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->x.p, xlist, sizeof(float) * c->npos, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->y.p, ylist, sizeof(float) * c->npos, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->Mtmp.p, measured, sizeof(float) * n2 * c->npos, hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_ingest_patterns(c->Mtmp.p, c->sqrtM.p, n, c->npos, c->p.diffraction_transpose != 0, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->Mtmp.release();
    c->uploaded = true;
    return MSET_OK;
}

int rotate_volume(mset_ctx* c, float* vol, const double* v1, const double* v2, const double* v3, const double* ang,
                  bool inverse) {
    double R[9];
    if (!rot_matrix(v1, v2, v3, ang[0], ang[1], ang[2], inverse, R)) return MSET_OK;   // |angles| <= 0.5
    if (!rot_size_supported(c->n) || !rot_size_supported(c->nz))
        return fail(MSET_EUNSUPPORTED, "rotation line length not built");
    Timed t(c, "rotate");
    const int r = rot_run(c->ws->rot, vol, c->n, c->n, c->nz, R, 11.0, c->stream);
    if (r != MSET_OK) return fail(r, rot_last_error());
    return MSET_OK;
}

}  // namespace

// ================================================================================================
struct mset_recon {
    int device = 0;
    hipStream_t stream = nullptr;
    std::shared_ptr<Workspace> ws;
    std::shared_ptr<ProfMap> prof;
    bool profiling = false;
    mset_params common{};
    int n = 0, nz = 0;
    double vec[9] = {0, 0, 1, 0, 1, 0, 1, 0, 0};
    DevBuf<float> rec, rec_snap;
    std::vector<float> prop_host;
    std::vector<mset_ctx*> tilts;
    std::vector<std::array<double, 3>> angles;
    std::vector<std::vector<float>> snap_probe, snap_x, snap_y;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_key[3] = {-1, -1, -1};
    // the error pass over every tilt in one launch: per-tilt table, work items, per-position errors
    DevBuf<TiltArgs> tilt_tab;
    DevBuf<int> items;
    DevBuf<double> err_all;
};

namespace {

void drop_graph(mset_recon* r) {
    if (r->exec) (void)hipGraphExecDestroy(r->exec);
    if (r->graph) (void)hipGraphDestroy(r->graph);
    r->exec = nullptr;
    r->graph = nullptr;
    r->graph_key[0] = r->graph_key[1] = r->graph_key[2] = -1;
}

int check_recon(mset_recon* r) {
    if (!r) return fail(MSET_EINVAL, "null reconstruction");
    HIPCHK(hipSetDevice(r->device));
    for (size_t j = 0; j < r->tilts.size(); ++j)
        if (!r->tilts[j] || !r->tilts[j]->uploaded)
            return fail(MSET_EINVAL, "tilt " + std::to_string(j) + " has no data (call mset_recon_set_tilt)");
    return MSET_OK;
}

// one update pass over every tilt (Run_MSET_recon.m:119-187), issue only
int issue_update_pass(mset_recon* r, int path, int batch, int positivity) {
    const size_t nv = (size_t)r->n * r->n * r->nz;
    for (size_t j = 0; j < r->tilts.size(); ++j) {
        mset_ctx* c = r->tilts[j];
        // STEP01: RVol = inv_rot(rec)
        HIPCHK(hipMemcpyAsync(c->V.p, r->rec.p, sizeof(float) * nv, hipMemcpyDeviceToDevice, r->stream));
        CHK(rotate_volume(c, c->V.p, r->vec, r->vec + 3, r->vec + 6, r->angles[j].data(), true));
        // STEP02-04
        if (path == MSET_PATH_PLAIN) CHK(issue_plain(c));
        else CHK(issue_batched(c, batch));
        // STEP05: rec = rot(RVol), positivity
        HIPCHK(hipMemcpyAsync(r->rec.p, c->V.p, sizeof(float) * nv, hipMemcpyDeviceToDevice, r->stream));
        CHK(rotate_volume(c, r->rec.p, r->vec, r->vec + 3, r->vec + 6, r->angles[j].data(), false));
        if (positivity) HIPCHK(launch_positivity(r->rec.p, nv, r->stream));
    }
    return MSET_OK;
}

}  // namespace

// ================================================================================================
extern "C" {

const char* mset_last_error(void) { return g_err.c_str(); }

int mset_create(mset_ctx** out, int device) {
    if (!out) return fail(MSET_EINVAL, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(MSET_EINVAL, "device index out of range");
    HIPCHK(hipSetDevice(device));
    return new_ctx(out, device, nullptr, nullptr, nullptr, nullptr);
}

int mset_destroy(mset_ctx* c) {
    free_ctx(c);
    return MSET_OK;
}

void* mset_stream(mset_ctx* c) { return c ? (void*)c->stream : nullptr; }

int mset_upload(mset_ctx* c, const mset_params* p, const float* RVol, const float* init_wave2D, const float* prop2D,
                const float* xlist, const float* ylist, const float* measured) {
    CHK(check_ctx(c));
    if (!RVol) return fail(MSET_EINVAL, "null input array");
    return upload_impl(c, p, RVol, init_wave2D, prop2D, xlist, ylist, measured);
}

int mset_run_resident(mset_ctx* c, int method, double* err_out) {
    CHK(check_uploaded(c));
    CHK(set_method(c, method));
    CHK(issue_plain(c));
    if (err_out) return read_errors(c, 0, c->npos, err_out);
    HIPCHK(hipStreamSynchronize(c->stream));
    return MSET_OK;
}

int mset_error_resident(mset_ctx* c, int method, double* err_out) {
    CHK(check_uploaded(c));
    CHK(set_method(c, method));
    if (!err_out) return fail(MSET_EINVAL, "null error output");
    CHK(issue_fwd(c, nullptr));
    return read_errors(c, 0, c->npos, err_out);
}

int mset_simulate(mset_ctx* c, float* intensity_out) {
    CHK(check_uploaded(c));
    CHK(set_method(c, MSET_METHOD_MSET));
    if (!intensity_out) return fail(MSET_EINVAL, "null output");
    DevBuf<float> I;
    CHK(I.ensure(plane(c) * c->npos, "intensities"));
    CHK(issue_fwd(c, I.p));
    HIPCHK(hipMemcpyAsync(intensity_out, I.p, sizeof(float) * plane(c) * c->npos, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MSET_OK;
}

int mset_download(mset_ctx* c, float* RVol, float* probe, float* xl, float* yl) {
    CHK(check_uploaded(c));
    const size_t n2 = plane(c);
    if (RVol) HIPCHK(hipMemcpyAsync(RVol, c->V.p, sizeof(float) * n2 * c->nz, hipMemcpyDeviceToHost, c->stream));
    if (probe) HIPCHK(hipMemcpyAsync(probe, c->probe.p, sizeof(float2) * n2, hipMemcpyDeviceToHost, c->stream));
    if (xl) HIPCHK(hipMemcpyAsync(xl, c->x.p, sizeof(float) * c->npos, hipMemcpyDeviceToHost, c->stream));
    if (yl) HIPCHK(hipMemcpyAsync(yl, c->y.p, sizeof(float) * c->npos, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MSET_OK;
}

int mset_set_state(mset_ctx* c, const float* RVol, const float* probe, const float* xl, const float* yl) {
    CHK(check_uploaded(c));
    const size_t n2 = plane(c);
    if (RVol) HIPCHK(hipMemcpyAsync(c->V.p, RVol, sizeof(float) * n2 * c->nz, hipMemcpyHostToDevice, c->stream));
    if (probe) HIPCHK(hipMemcpyAsync(c->probe.p, probe, sizeof(float2) * n2, hipMemcpyHostToDevice, c->stream));
    if (xl) HIPCHK(hipMemcpyAsync(c->x.p, xl, sizeof(float) * c->npos, hipMemcpyHostToDevice, c->stream));
    if (yl) HIPCHK(hipMemcpyAsync(c->y.p, yl, sizeof(float) * c->npos, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MSET_OK;
}

int mset_run(mset_ctx* c, int method, float* RVol_out, float* probe_out, float* xl_out, float* yl_out,
             float* err_out) {
    double e = 0.0;
    CHK(mset_run_resident(c, method, err_out ? &e : nullptr));
    CHK(mset_download(c, RVol_out, probe_out, xl_out, yl_out));
    if (err_out) *err_out = (float)e;
    return MSET_OK;
}

int mset_error(mset_ctx* c, int method, float* err_out) {
    if (!err_out) return fail(MSET_EINVAL, "null error output");
    double e = 0.0;
    CHK(mset_error_resident(c, method, &e));
    *err_out = (float)e;
    return MSET_OK;
}

int mset_grad(mset_ctx* c, int first, int count, float* dV, float* dprobe, float* dpos, double* err_out) {
    CHK(check_uploaded(c));
    if (first < 0 || count <= 0 || first + count > c->npos) return fail(MSET_EINVAL, "position range out of bounds");
    CHK(set_method(c, MSET_METHOD_MSET));
    CHK(issue_grad(c, first, count, false, WANT_PROBE | WANT_POS));
    const size_t n2 = plane(c);
    if (dV) HIPCHK(hipMemcpyAsync(dV, c->ws->G.p, sizeof(float) * n2 * c->L, hipMemcpyDeviceToHost, c->stream));
    if (dprobe) HIPCHK(hipMemcpyAsync(dprobe, c->ws->PG.p, sizeof(float2) * n2, hipMemcpyDeviceToHost, c->stream));
    if (dpos) HIPCHK(hipMemcpyAsync(dpos, c->dpos.p + 2 * (size_t)first, sizeof(float) * 2 * count,
                                    hipMemcpyDeviceToHost, c->stream));
    if (err_out) return read_errors(c, first, count, err_out);
    HIPCHK(hipStreamSynchronize(c->stream));
    return MSET_OK;
}

int mset_grad_noalloc(mset_ctx* c, int first, int count, float* dV, float* dprobe, float* dpos, double* err_pos) {
    CHK(check_uploaded(c));
    if (first < 0 || count <= 0 || first + count > c->npos) return fail(MSET_EINVAL, "position range out of bounds");
    if (!dV || !dprobe || !dpos || !err_pos) return fail(MSET_EINVAL, "null output (device pointers required)");
    CHK(set_method(c, MSET_METHOD_MSET));
    unsigned char* base = scratch_on(c->device);
    if (!base) return fail(MSET_EDEVICE, "scratch array not resolved on this device");
    const size_t n2 = plane(c), L = (size_t)c->L;
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    // fixed part: transmission planes and fftn(probe).  Per gradient slab: L object planes and a probe plane.
    // Per resident position (slot): stored waves (L + 2 planes) and, at N = 256, the engine's two plane
    // buffers.  Four blocks with the kernels' plane strides, each starting on a 256-byte boundary.
    // the separable kernels (as mset_grad picks them) keep the transmission in two layouts
    const bool sep = sep_fwd(c) && c->p.grad_mode == GRAD_REF;
    const size_t b_trans = up(L * n2 * sizeof(float2)), b_pf = up(n2 * sizeof(float2));
    const size_t s_g = L * n2 * sizeof(float), s_ps = n2 * sizeof(float2);
    const size_t s_save = (L + 2) * n2 * sizeof(float2), s_pl = plane_in_global(c->n) ? 2 * n2 * sizeof(float2) : 0;
    const size_t fixed = b_trans + b_pf + 4 * 256;
    auto need = [&](int S, int G) { return fixed + (size_t)S * (s_g + s_ps) + (size_t)G * (s_save + s_pl); };
    // mset_grad's slabs and grid when they fit (then the sums are bitwise mset_grad's), else as many resident
    // positions as fit, one slab each
    const int shape = shape_for(c, WAVES_STORE);
    int S = std::min(grad_slabs(c, shape), count);
    int G = launch_grid(batch_slots(c, shape), count, S);
    while (G > 1 && need(S, G) > MSET_NOALLOC_SCRATCH_BYTES) G = launch_grid(G - 1, count, S);
    if (need(S, G) > MSET_NOALLOC_SCRATCH_BYTES) {
        S = G = (int)std::min<size_t>((size_t)count, (MSET_NOALLOC_SCRATCH_BYTES - std::min(fixed, (size_t)MSET_NOALLOC_SCRATCH_BYTES)) /
                                                         (s_g + s_ps + s_save + s_pl));
        if (S < 1)
            return fail(MSET_EINVAL, "mset_grad_noalloc: this problem needs " + std::to_string(need(1, 1)) +
                                         " bytes of scratch for one resident position; the maximum is " +
                                         std::to_string((size_t)MSET_NOALLOC_SCRATCH_BYTES));
    }
    unsigned char* q = base;
    float2* trans = (float2*)q;  q += b_trans;
    float2* probeF = (float2*)q; q += b_pf;
    float* gslab = (float*)q;    q += up((size_t)S * s_g);
    float2* pslab = (float2*)q;  q += up((size_t)S * s_ps);
    float2* save = (float2*)q;   q += up((size_t)G * s_save);
    float2* planes_ = s_pl ? (float2*)q : nullptr;
    if (sep)
        HIPCHK(launch_transmission_sep(c->V.p, (float*)trans, (float*)trans + L * n2, c->n, c->nz, binning(c), c->p.sigma,
                                       c->p.potential_pixelsize, c->stream));
    else
        HIPCHK(launch_transmission(c->V.p, trans, c->n, c->nz, binning(c), 0, c->p.sigma, c->p.potential_pixelsize,
                                   c->stream));
    HIPCHK(launch_plane_fft(c->n, c->probe.p, probeF, c->tw.p, 0, 1.f, 0.f, planes_, c->stream));
    PosArgs A = base_args(c);
    A.trans = trans;
    A.probeF = probeF;
    A.save = save;
    A.waves = WAVES_STORE;
    A.gslab = gslab;
    A.pslab = pslab;
    A.planes = planes_;
    A.dpos = dpos;
    A.err_pos = err_pos;
    A.obase = first;
    A.first = first;
    A.count = count;
    A.want = WANT_PROBE | WANT_POS;
    A.nslab = S;
    A.sep = sep ? 1 : 0;
    A.tph = (const float*)trans;
    A.tph_b = (const float*)trans + L * n2;
    HIPCHK(launch_batch(c->n, A, G, batch_shape(c, shape, sep, WAVES_STORE, count), c->stream));
    HIPCHK(launch_reduce_slabs(gslab, S, L * n2, dV, c->stream));
    if (sep) HIPCHK(launch_untranspose_odd(dV, c->n, c->L, c->stream));
    HIPCHK(launch_reduce_cslabs(pslab, S, n2, pslab, c->stream));
    HIPCHK(launch_plane_fft(c->n, pslab, (float2*)dprobe, c->tw.p, 1, 1.f, 0.f, planes_, c->stream));
    return MSET_OK;
}

int mset_waves(mset_ctx* c, int k, float* waves_out) {
    CHK(check_uploaded(c));
    if (k < 0 || k >= c->npos || !waves_out) return fail(MSET_EINVAL, "bad position index or null output");
    CHK(set_method(c, MSET_METHOD_MSET));
    CHK(ensure_batch_buffers(c));
    CHK(prepare_fixed(c));
    // every plane's wave is read back: the store layout, in slot 0
    CHK(c->ws->save.ensure((size_t)(c->L + 2) * plane(c), "save waves"));
    PosArgs A = base_args(c);
    A.first = k;
    A.count = 1;
    A.want = 0;
    A.waves = WAVES_STORE;
    A.nslab = 1;
    HIPCHK(launch_batch(c->n, A, 1, 0, c->stream));
    HIPCHK(hipMemcpyAsync(waves_out, c->ws->save.p, sizeof(float2) * (size_t)(c->L + 1) * plane(c),
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MSET_OK;
}

int mset_run_batched(mset_ctx* c, int method, int batch, double* err_out) {
    CHK(check_uploaded(c));
    CHK(set_method(c, method));
    if (batch <= 0) return fail(MSET_EINVAL, "batch must be positive");
    CHK(issue_batched(c, batch));
    if (err_out) return read_errors(c, 0, c->npos, err_out);
    HIPCHK(hipStreamSynchronize(c->stream));
    return MSET_OK;
}

int mset_set_exit_waves(mset_ctx* c, int mode) {
    CHK(check_ctx(c));
    if (mode != MSET_EXIT_WAVES_AUTO && mode != MSET_EXIT_WAVES_STORE && mode != MSET_EXIT_WAVES_RECOMPUTE &&
        mode != MSET_EXIT_WAVES_CHECKPOINT)
        return fail(MSET_EINVAL, "bad exit-wave mode");
    if (mode != c->ws->exit_waves) {
        // the next batched call sizes the wave scratch for the new variant
        HIPCHK(hipStreamSynchronize(c->stream));
        c->ws->save.release();
        c->ws->exit_waves = mode;
    }
    return MSET_OK;
}

int mset_set_slot_limit(mset_ctx* c, int slots) {
    CHK(check_ctx(c));
    if (slots < 0) return fail(MSET_EINVAL, "slot limit must be >= 0");
    c->ws->slot_limit = slots;
    return MSET_OK;
}

int mset_set_engine(mset_ctx* c, int engine) {
    CHK(check_ctx(c));
    if (engine & ~(MSET_ENGINE_GENERAL | MSET_ENGINE_ALT_SHAPE)) return fail(MSET_EINVAL, "bad engine");
    c->ws->engine = engine;
    return MSET_OK;
}

int mset_exit_waves(mset_ctx* c, int* resolved, size_t* bytes_per_slot) {
    CHK(check_ctx(c));
    if (resolved)
        *resolved = c->ws->waves == WAVES_RECOMPUTE    ? MSET_EXIT_WAVES_RECOMPUTE
                    : c->ws->waves == WAVES_CHECKPOINT ? MSET_EXIT_WAVES_CHECKPOINT
                                                       : MSET_EXIT_WAVES_STORE;
    if (bytes_per_slot) *bytes_per_slot = c->ws->slot_bytes;
    return MSET_OK;
}

int mset_rotate(mset_ctx* c, const float* in, float* out, int nx, int ny, int nz, const double* R, double step) {
    CHK(check_ctx(c));
    if (!in || !out || !R) return fail(MSET_EINVAL, "null argument");
    if (nx <= 0 || ny <= 0 || nz <= 0) return fail(MSET_EINVAL, "sizes must be positive");
    if (!rot_size_supported(nx) || !rot_size_supported(ny) || !rot_size_supported(nz))
        return fail(MSET_EUNSUPPORTED, "rotation line length not built");
    const size_t nv = (size_t)nx * ny * nz;
    DevBuf<float> buf;
    CHK(buf.ensure(nv, "rotation volume"));
    HIPCHK(hipMemcpyAsync(buf.p, in, sizeof(float) * nv, hipMemcpyHostToDevice, c->stream));
    {
        Timed t(c, "rotate");
        const int r = rot_run(c->ws->rot, buf.p, nx, ny, nz, R, step, c->stream);
        if (r != MSET_OK) return fail(r, rot_last_error());
    }
    HIPCHK(hipMemcpyAsync(out, buf.p, sizeof(float) * nv, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MSET_OK;
}

int mset_device_memory(int device, size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return fail(MSET_EINVAL, "null output");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(MSET_EINVAL, "device index out of range");
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipMemGetInfo(free_bytes, total_bytes));
    return MSET_OK;
}

int mset_host_alloc(size_t bytes, void** ptr) {
    if (!ptr) return fail(MSET_EINVAL, "null output");
    *ptr = nullptr;
    if (bytes == 0) return fail(MSET_EINVAL, "zero-byte host buffer");
    hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(MSET_ENOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    return MSET_OK;
}

int mset_host_free(void* ptr) {
    if (!ptr) return MSET_OK;
    HIPCHK(hipHostFree(ptr));
    return MSET_OK;
}

int mset_batch_slots(int device, int n, int* slots) {
    if (!slots) return fail(MSET_EINVAL, "null output");
    if (!plane_size_supported(n)) return fail(MSET_EUNSUPPORTED, "plane size not built");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(MSET_EINVAL, "device index out of range");
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    int per_cu = 1;
    HIPCHK(occupancy_batch(n, 0, &per_cu));
    *slots = prop.multiProcessorCount * (per_cu < 1 ? 1 : per_cu);
    return MSET_OK;
}

int mset_tv(mset_ctx* c, const float* in, float* out, int nx, int ny, int nz, float lambda, int n_iter,
            double* tv_error) {
    CHK(check_ctx(c));
    if (!in || !out) return fail(MSET_EINVAL, "null argument");
    if (nx <= 0 || ny <= 0 || nz <= 0) return fail(MSET_EINVAL, "sizes must be positive");
    if (!(lambda > 0.f) || n_iter < 0) return fail(MSET_EINVAL, "TV needs lambda > 0 and n_iter >= 0");
    const long long nv = (long long)nx * ny * nz;
    DevBuf<float> buf;
    CHK(buf.ensure((size_t)nv, "TV volume"));
    if (tv_ensure(c->ws->tv, nv) != 0) return fail(MSET_ENOMEM, "hipMalloc: TV work");
    HIPCHK(hipMemcpyAsync(buf.p, in, sizeof(float) * nv, hipMemcpyHostToDevice, c->stream));
    HIPCHK(tv_run(c->ws->tv, buf.p, buf.p, nx, ny, nz, lambda, n_iter, tv_error, c->stream));
    HIPCHK(hipMemcpyAsync(out, buf.p, sizeof(float) * nv, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MSET_OK;
}

int mset_rotate_resident(mset_ctx* c, const double* v1, const double* v2, const double* v3, double phi, double theta,
                         double psi, int inverse) {
    CHK(check_uploaded(c));
    if (!v1 || !v2 || !v3) return fail(MSET_EINVAL, "null rotation axis");
    const double ang[3] = {phi, theta, psi};
    CHK(rotate_volume(c, c->V.p, v1, v2, v3, ang, inverse != 0));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MSET_OK;
}

int mset_profile_enable(mset_ctx* c, int enable) {
    CHK(check_ctx(c));
    *c->profiling = enable != 0;
    return MSET_OK;
}

int mset_profile_read(mset_ctx* c, const char* name, int* launches, double* total_ms) {
    CHK(check_ctx(c));
    if (!name) return fail(MSET_EINVAL, "null name");
    return prof_read(*c->prof, c->stream, name, launches, total_ms);
}

int mset_synchronize(mset_ctx* c) {
    CHK(check_ctx(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MSET_OK;
}

// ---------------------------------------------------------------------------------------- recon
int mset_recon_create(mset_recon** out, int device, const mset_params* common, int n_tilts, const double* vecs,
                      const float* rec, const float* prop2D) {
    if (!out || !common || !rec || !prop2D) return fail(MSET_EINVAL, "null argument");
    *out = nullptr;
    if (n_tilts <= 0) return fail(MSET_EINVAL, "n_tilts must be positive");
    if (common->nx != common->ny || common->nx <= 0 || common->nz <= 0)
        return fail(MSET_EINVAL, "volume must be square in x, y with positive sizes");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(MSET_EINVAL, "device index out of range");
    HIPCHK(hipSetDevice(device));
    auto r = std::make_unique<mset_recon>();
    r->device = device;
    r->common = *common;
    r->n = common->nx;
    r->nz = common->nz;
    if (vecs) std::memcpy(r->vec, vecs, sizeof(r->vec));
    HIPCHK(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    r->ws = std::make_shared<Workspace>();
    r->prof = std::make_shared<ProfMap>();
    const size_t nv = (size_t)r->n * r->n * r->nz;
    CHK(r->rec.ensure(nv, "reconstruction"));
    HIPCHK(hipMemcpyAsync(r->rec.p, rec, sizeof(float) * nv, hipMemcpyHostToDevice, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    r->prop_host.assign(prop2D, prop2D + 2 * (size_t)r->n * r->n);
    r->tilts.assign(n_tilts, nullptr);
    r->angles.assign(n_tilts, {0, 0, 0});
    *out = r.release();
    return MSET_OK;
}

int mset_recon_destroy(mset_recon* r) {
    if (!r) return MSET_OK;
    (void)hipSetDevice(r->device);
    (void)hipStreamSynchronize(r->stream);
    drop_graph(r);
    for (mset_ctx* c : r->tilts) free_ctx(c);
    prof_free(*r->prof);
    r->ws.reset();
    (void)hipStreamDestroy(r->stream);
    delete r;
    return MSET_OK;
}

int mset_recon_set_tilt(mset_recon* r, int j, const double* angles, const float* probe, int n_scan_x, int n_scan_y,
                        int num_scan_pos, const float* xlist, const float* ylist, const float* measured) {
    if (!r) return fail(MSET_EINVAL, "null reconstruction");
    HIPCHK(hipSetDevice(r->device));
    if (j < 0 || j >= (int)r->tilts.size()) return fail(MSET_EINVAL, "tilt index out of range");
    if (!angles) return fail(MSET_EINVAL, "null angles");
    drop_graph(r);
    if (!r->tilts[j]) CHK(new_ctx(&r->tilts[j], r->device, r->stream, r->ws, &r->profiling, r->prof));
    mset_params p = r->common;
    p.n_scan_x = n_scan_x;
    p.n_scan_y = n_scan_y;
    p.num_scan_pos = num_scan_pos;
    CHK(upload_impl(r->tilts[j], &p, nullptr, probe, r->prop_host.data(), xlist, ylist, measured));
    r->angles[j] = {angles[0], angles[1], angles[2]};
    return MSET_OK;
}

int mset_recon_set_steps(mset_recon* r, const float* steps) {
    if (!r || !steps) return fail(MSET_EINVAL, "null argument");
    for (int i = 0; i < 3; ++i) r->common.step_size[i] = steps[i];
    for (mset_ctx* c : r->tilts)
        if (c)
            for (int i = 0; i < 3; ++i) c->p.step_size[i] = steps[i];
    drop_graph(r);
    return MSET_OK;
}

int mset_recon_iterate(mset_recon* r, int path, int batch, int use_graph, int positivity, double* err_out) {
    CHK(check_recon(r));
    if (path != MSET_PATH_PLAIN && path != MSET_PATH_BATCHED) return fail(MSET_EINVAL, "bad path");
    if (path == MSET_PATH_BATCHED && batch <= 0) return fail(MSET_EINVAL, "batch must be positive");
    if (path == MSET_PATH_PLAIN) batch = 1;
    if (use_graph && !r->profiling) {
        const bool same = r->exec && r->graph_key[0] == path && r->graph_key[1] == batch && r->graph_key[2] == positivity;
        if (!same) {
            drop_graph(r);
            // eager pass first: every allocation and twiddle upload happens outside the capture
            CHK(issue_update_pass(r, path, batch, positivity));
            HIPCHK(hipStreamSynchronize(r->stream));
            HIPCHK(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
            const int rc = issue_update_pass(r, path, batch, positivity);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(r->stream, &g);
            if (rc != MSET_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            HIPCHK(ec);
            r->graph = g;
            HIPCHK(hipGraphInstantiate(&r->exec, r->graph, nullptr, nullptr, 0));
            r->graph_key[0] = path;
            r->graph_key[1] = batch;
            r->graph_key[2] = positivity;
            // the eager pass above was this call's iteration
        } else {
            HIPCHK(hipGraphLaunch(r->exec, r->stream));
        }
    } else {
        CHK(issue_update_pass(r, path, batch, positivity));
    }
    if (err_out) {
        HIPCHK(hipStreamSynchronize(r->stream));
        double s = 0.0;
        for (mset_ctx* c : r->tilts) {
            double e = 0.0;
            CHK(read_errors(c, 0, c->npos, &e));
            s += e;
        }
        *err_out = s / (double)r->tilts.size();
    }
    HIPCHK(hipStreamSynchronize(r->stream));
    return MSET_OK;
}

// The error pass of Run_MSET_recon.m:189-239 (run_error_mset, mset.cu:557-696) over every tilt: each tilt's
// rotated volume, transmission planes and probe spectrum are prepared, then the positions of all tilts are
// one persistent launch (no per-tilt launch tails or host synchronisations), and the per-position errors
// come back in one copy.
int mset_recon_error_pass(mset_recon* r, double* errs) {
    CHK(check_recon(r));
    if (!errs) return fail(MSET_EINVAL, "null output");
    const int T = (int)r->tilts.size();
    if (T >= (1 << (31 - ITEM_TILT_SHIFT))) return fail(MSET_EINVAL, "too many tilts for one error-pass launch");
    const size_t nv = (size_t)r->n * r->n * r->nz;
    mset_ctx* c0 = r->tilts[0];
    std::vector<size_t> off(T + 1, 0);
    for (int j = 0; j < T; ++j) {
        if (r->tilts[j]->npos > ITEM_POS_MASK) return fail(MSET_EINVAL, "too many positions in one tilt");
        off[j + 1] = off[j] + r->tilts[j]->npos;
    }
    const size_t tot = off[T];
    CHK(r->err_all.ensure(tot, "error-pass errors"));
    CHK(r->items.ensure(tot, "error-pass work items"));
    CHK(r->tilt_tab.ensure(T, "error-pass tilt table"));
    int grid = std::max(1, device_slots(c0, 0));
    if (r->ws->slot_limit > 0) grid = std::min(grid, r->ws->slot_limit);
    grid = (int)std::min<size_t>((size_t)grid, tot);
    if (plane_in_global(r->n)) {
        const size_t need = (size_t)grid * 2 * plane(c0);
        if (r->ws->planes.n < need) drop_graph(r);       // a captured update pass holds the current buffer
        CHK(r->ws->planes.ensure(need, "plane buffers"));
    }
    std::vector<TiltArgs> tab(T);
    std::vector<int> items(tot);
    const bool sep = sep_fwd(c0);
    for (int j = 0; j < T; ++j) {
        mset_ctx* c = r->tilts[j];
        CHK(set_method(c, MSET_METHOD_MSET));
        HIPCHK(hipMemcpyAsync(c->V.p, r->rec.p, sizeof(float) * nv, hipMemcpyDeviceToDevice, r->stream));
        CHK(rotate_volume(c, c->V.p, r->vec, r->vec + 3, r->vec + 6, r->angles[j].data(), true));
        CHK(c->mt_trans.ensure((size_t)c->L * plane(c), "tilt transmission"));
        CHK(c->mt_probeF.ensure(plane(c), "tilt probe spectrum"));
        if (sep)
            HIPCHK(launch_transmission_sep(c->V.p, (float*)c->mt_trans.p, nullptr, c->n, c->nz, binning(c), c->p.sigma,
                                           c->p.potential_pixelsize, r->stream));
        else
            HIPCHK(launch_transmission(c->V.p, c->mt_trans.p, c->n, c->nz, binning(c), 0, c->p.sigma,
                                       c->p.potential_pixelsize, r->stream));
        HIPCHK(launch_plane_fft(c->n, c->probe.p, c->mt_probeF.p, c->tw.p, 0, 1.f, 0.f, r->ws->planes.p, r->stream));
        tab[j] = TiltArgs{c->mt_trans.p, (const float*)c->mt_trans.p, c->mt_probeF.p, c->sqrtM.p, c->x.p, c->y.p,
                          r->err_all.p + off[j]};
        for (int k = 0; k < c->npos; ++k) items[off[j] + k] = (j << ITEM_TILT_SHIFT) | k;   // tilt-major
    }
    HIPCHK(hipMemcpyAsync(r->tilt_tab.p, tab.data(), sizeof(TiltArgs) * T, hipMemcpyHostToDevice, r->stream));
    HIPCHK(hipMemcpyAsync(r->items.p, items.data(), sizeof(int) * tot, hipMemcpyHostToDevice, r->stream));
    PosArgs A = base_args(c0);
    A.tilts = r->tilt_tab.p;
    A.order = r->items.p;
    A.first = 0;
    A.count = (int)tot;
    A.planes = plane_in_global(r->n) ? r->ws->planes.p : nullptr;
    A.sep = sep ? 1 : 0;
    {
        Timed t(c0, "position_fwd");
        HIPCHK(launch_fwd_tilts(r->n, A, grid, r->stream));
    }
    std::vector<double> e(tot);
    HIPCHK(hipMemcpyAsync(e.data(), r->err_all.p, sizeof(double) * tot, hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    for (int j = 0; j < T; ++j)
        errs[j] = host_error_sum(r->tilts[j], std::vector<double>(e.begin() + off[j], e.begin() + off[j + 1]));
    return MSET_OK;
}

// rec = sf_max * TV_method(rec / sf_max, lambda, n_iter), sf_max = max(rec(:)) (Run_MSET_recon.m:266-269)
int mset_recon_tv(mset_recon* r, float lambda, int n_iter, double* tv_error) {
    CHK(check_recon(r));
    if (!(lambda > 0.f) || n_iter < 0) return fail(MSET_EINVAL, "TV needs lambda > 0 and n_iter >= 0");
    const long long nv = (long long)r->n * r->n * r->nz;
    if (tv_ensure(r->ws->tv, nv) != 0) return fail(MSET_ENOMEM, "hipMalloc: TV work");
    float m = 0.f;
    HIPCHK(volume_max(r->ws->tv, r->rec.p, nv, &m, r->stream));
    if (!(m != 0.f) || !std::isfinite(m))
        return fail(MSET_EINVAL, "TV normalises by max(rec): the maximum is zero or not finite");
    drop_graph(r);
    HIPCHK(volume_scale(r->rec.p, nv, m, 0, r->stream));
    HIPCHK(tv_run(r->ws->tv, r->rec.p, r->rec.p, r->n, r->n, r->nz, lambda, n_iter, tv_error, r->stream));
    HIPCHK(volume_scale(r->rec.p, nv, m, 1, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return MSET_OK;
}

int mset_recon_get(mset_recon* r, float* rec, float* probes, float* xlists, float* ylists) {
    CHK(check_recon(r));
    const size_t n2 = (size_t)r->n * r->n;
    if (rec) HIPCHK(hipMemcpyAsync(rec, r->rec.p, sizeof(float) * n2 * r->nz, hipMemcpyDeviceToHost, r->stream));
    size_t off = 0;
    for (size_t j = 0; j < r->tilts.size(); ++j) {
        mset_ctx* c = r->tilts[j];
        if (probes)
            HIPCHK(hipMemcpyAsync(probes + 2 * n2 * j, c->probe.p, sizeof(float2) * n2, hipMemcpyDeviceToHost, r->stream));
        if (xlists) HIPCHK(hipMemcpyAsync(xlists + off, c->x.p, sizeof(float) * c->npos, hipMemcpyDeviceToHost, r->stream));
        if (ylists) HIPCHK(hipMemcpyAsync(ylists + off, c->y.p, sizeof(float) * c->npos, hipMemcpyDeviceToHost, r->stream));
        off += c->npos;
    }
    HIPCHK(hipStreamSynchronize(r->stream));
    return MSET_OK;
}

int mset_recon_set_exit_waves(mset_recon* r, int mode) {
    if (!r) return fail(MSET_EINVAL, "null reconstruction");
    HIPCHK(hipSetDevice(r->device));
    if (mode != MSET_EXIT_WAVES_AUTO && mode != MSET_EXIT_WAVES_STORE && mode != MSET_EXIT_WAVES_RECOMPUTE &&
        mode != MSET_EXIT_WAVES_CHECKPOINT)
        return fail(MSET_EINVAL, "bad exit-wave mode");
    if (mode != r->ws->exit_waves) {
        HIPCHK(hipStreamSynchronize(r->stream));
        drop_graph(r);   // the captured pass holds the old scratch and kernel variant
        r->ws->save.release();
        r->ws->exit_waves = mode;
    }
    return MSET_OK;
}

int mset_recon_set_slot_limit(mset_recon* r, int slots) {
    if (!r) return fail(MSET_EINVAL, "null reconstruction");
    if (slots < 0) return fail(MSET_EINVAL, "slot limit must be >= 0");
    HIPCHK(hipSetDevice(r->device));
    if (slots != r->ws->slot_limit) {
        HIPCHK(hipStreamSynchronize(r->stream));
        drop_graph(r);   // the captured pass holds the old grid
        r->ws->slot_limit = slots;
    }
    return MSET_OK;
}

int mset_recon_set_engine(mset_recon* r, int engine) {
    if (!r) return fail(MSET_EINVAL, "null reconstruction");
    if (engine & ~(MSET_ENGINE_GENERAL | MSET_ENGINE_ALT_SHAPE)) return fail(MSET_EINVAL, "bad engine");
    HIPCHK(hipSetDevice(r->device));
    if (engine != r->ws->engine) {
        HIPCHK(hipStreamSynchronize(r->stream));
        drop_graph(r);   // the captured pass holds the other routine's kernels and layouts
        r->ws->engine = engine;
    }
    return MSET_OK;
}

int mset_recon_snapshot(mset_recon* r) {
    CHK(check_recon(r));
    const size_t n2 = (size_t)r->n * r->n, nv = n2 * r->nz;
    CHK(r->rec_snap.ensure(nv, "snapshot"));
    HIPCHK(hipMemcpyAsync(r->rec_snap.p, r->rec.p, sizeof(float) * nv, hipMemcpyDeviceToDevice, r->stream));
    r->snap_probe.resize(r->tilts.size());
    r->snap_x.resize(r->tilts.size());
    r->snap_y.resize(r->tilts.size());
    for (size_t j = 0; j < r->tilts.size(); ++j) {
        mset_ctx* c = r->tilts[j];
        r->snap_probe[j].resize(2 * n2);
        r->snap_x[j].resize(c->npos);
        r->snap_y[j].resize(c->npos);
        HIPCHK(hipMemcpyAsync(r->snap_probe[j].data(), c->probe.p, sizeof(float2) * n2, hipMemcpyDeviceToHost, r->stream));
        HIPCHK(hipMemcpyAsync(r->snap_x[j].data(), c->x.p, sizeof(float) * c->npos, hipMemcpyDeviceToHost, r->stream));
        HIPCHK(hipMemcpyAsync(r->snap_y[j].data(), c->y.p, sizeof(float) * c->npos, hipMemcpyDeviceToHost, r->stream));
    }
    HIPCHK(hipStreamSynchronize(r->stream));
    return MSET_OK;
}

int mset_recon_restore(mset_recon* r) {
    CHK(check_recon(r));
    if (!r->rec_snap.p) return fail(MSET_EINVAL, "no snapshot taken");
    const size_t n2 = (size_t)r->n * r->n, nv = n2 * r->nz;
    HIPCHK(hipMemcpyAsync(r->rec.p, r->rec_snap.p, sizeof(float) * nv, hipMemcpyDeviceToDevice, r->stream));
    for (size_t j = 0; j < r->tilts.size(); ++j) {
        mset_ctx* c = r->tilts[j];
        HIPCHK(hipMemcpyAsync(c->probe.p, r->snap_probe[j].data(), sizeof(float2) * n2, hipMemcpyHostToDevice, r->stream));
        HIPCHK(hipMemcpyAsync(c->x.p, r->snap_x[j].data(), sizeof(float) * c->npos, hipMemcpyHostToDevice, r->stream));
        HIPCHK(hipMemcpyAsync(c->y.p, r->snap_y[j].data(), sizeof(float) * c->npos, hipMemcpyHostToDevice, r->stream));
    }
    HIPCHK(hipStreamSynchronize(r->stream));
    return MSET_OK;
}

int mset_recon_profile_enable(mset_recon* r, int enable) {
    if (!r) return fail(MSET_EINVAL, "null reconstruction");
    r->profiling = enable != 0;
    if (r->profiling) drop_graph(r);
    return MSET_OK;
}

int mset_recon_profile_read(mset_recon* r, const char* name, int* launches, double* total_ms) {
    if (!r || !name) return fail(MSET_EINVAL, "null argument");
    HIPCHK(hipSetDevice(r->device));
    return prof_read(*r->prof, r->stream, name, launches, total_ms);
}

void* mset_recon_stream(mset_recon* r) { return r ? (void*)r->stream : nullptr; }

}  // extern "C"
