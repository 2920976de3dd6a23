// fft_lds.hpp -- LDS-resident 2D FFT of one N x N complex plane for CDNA4 (gfx950).
//
// The plane lives in LDS as float2 s[N * S] with row pitch S = N + 1: element (y, x) (x fastest, the
// MATLAB dim-1 axis) is s[y * S + x].  A 2D transform is a pass of 1D transforms along y (columns)
// and a pass along x (rows); each pass is a sequence of Stockham autosort stages with radices
// Plan<N>::R.  In one stage every thread loads the R points of its butterflies into registers, the
// workgroup synchronises, every thread applies twiddles and the R-point DFT in registers and writes
// the results back in place, and the workgroup synchronises again.  Data is in natural order after
// every complete 1D pass, so there is no bit-reversal pass.
//
// LDS banks: in both passes consecutive lanes own consecutive lines (x in the column pass, y in the row
// pass) at the same position along the line.  A column-pass access is therefore a contiguous run of
// float2, and a row-pass access strides by the odd pitch S, which spreads 16 consecutive lanes over all
// banks: neither pass has bank conflicts, whatever the stage stride.
//
// Element-wise work is fused into the transform instead of costing LDS sweeps of its own: `pre(p, v)`
// is applied to every element p as the first pass loads it (with LOADG the functor supplies the value
// from global memory and LDS is not read) and `post(p, v)` as the last pass stores it.  A forward
// transform (FWD_CR) runs columns then rows, an inverse (INV_RC) rows then columns.  In a column pass
// consecutive lanes see consecutive x, so functors touching row-major [y][x] global arrays are
// coalesced; in a row pass consecutive lanes see consecutive y, so functors there should touch arrays
// stored transposed, [x][y].  Callers choose layouts accordingly (the propagator, for instance, is only
// ever applied in the row pass that ends a forward transform and is kept transposed).
//
// Forward transforms use exp(-2 pi i / N) (MATLAB fftn), inverse ones exp(+2 pi i / N) without the
// 1/N^2 factor (callers fold it into a post functor).  Twiddles come from a table tw[k] = W_N^k,
// k < N, computed on the host in double precision.
#pragma once
#include <hip/hip_runtime.h>

namespace mset {

// ------------------------------------------------------------------ complex helpers
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV> __device__ __forceinline__ float2 mul_mi(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ------------------------------------------------------------------ radix plans
template <int N> struct Plan;
template <> struct Plan<16>  { static constexpr int S = 2; static constexpr int R[2] = {4, 4}; };
template <> struct Plan<24>  { static constexpr int S = 2; static constexpr int R[2] = {8, 3}; };
template <> struct Plan<32>  { static constexpr int S = 2; static constexpr int R[2] = {8, 4}; };
template <> struct Plan<48>  { static constexpr int S = 2; static constexpr int R[2] = {16, 3}; };
template <> struct Plan<64>  { static constexpr int S = 2; static constexpr int R[2] = {8, 8}; };
template <> struct Plan<96>  { static constexpr int S = 3; static constexpr int R[3] = {8, 4, 3}; };
template <> struct Plan<128> { static constexpr int S = 2; static constexpr int R[2] = {16, 8}; };
template <> struct Plan<256> { static constexpr int S = 2; static constexpr int R[2] = {16, 16}; };

// LDS row pitch of an N x N plane
template <int N> struct Pitch { static constexpr int S = N + 1; };

// ------------------------------------------------------------------ small DFTs in registers
// W_64^j = cos(2 pi j/64) - i sin(2 pi j/64) for j = 0..31; W_R^k = W_64^(k 64/R)
__device__ __forceinline__ float2 w64(int j, bool inv) {
    constexpr float c[32] = {
        1.0f, 0.99518472667219693f, 0.98078528040323043f, 0.95694033573220882f, 0.92387953251128674f,
        0.88192126434835505f, 0.83146961230254524f, 0.77301045336273699f, 0.70710678118654757f,
        0.63439328416364549f, 0.55557023301960229f, 0.47139673682599781f, 0.38268343236508984f,
        0.29028467725446233f, 0.19509032201612833f, 0.09801714032956077f, 0.0f, -0.098017140329560645f,
        -0.19509032201612819f, -0.29028467725446216f, -0.38268343236508973f, -0.4713967368259977f,
        -0.55557023301960196f, -0.63439328416364538f, -0.70710678118654746f, -0.77301045336273699f,
        -0.83146961230254535f, -0.88192126434835494f, -0.92387953251128674f, -0.95694033573220882f,
        -0.98078528040323043f, -0.99518472667219682f};
    constexpr float s[32] = {
        0.0f, 0.098017140329560604f, 0.19509032201612825f, 0.29028467725446233f, 0.38268343236508978f,
        0.47139673682599764f, 0.55557023301960218f, 0.63439328416364549f, 0.70710678118654746f,
        0.77301045336273699f, 0.83146961230254524f, 0.88192126434835494f, 0.92387953251128674f,
        0.95694033573220894f, 0.98078528040323043f, 0.99518472667219682f, 1.0f, 0.99518472667219693f,
        0.98078528040323043f, 0.95694033573220894f, 0.92387953251128674f, 0.88192126434835505f,
        0.83146961230254546f, 0.7730104533627371f, 0.70710678118654757f, 0.63439328416364549f,
        0.55557023301960218f, 0.47139673682599786f, 0.38268343236508989f, 0.29028467725446239f,
        0.19509032201612861f, 0.098017140329560826f};
    return make_float2(c[j], inv ? s[j] : -s[j]);
}

template <int R, bool INV> struct DFT;

template <bool INV> struct DFT<2, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV> struct DFT<4, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        const float2 b0 = cadd(v[1], v[3]), b1 = mul_mi<INV>(csub(v[1], v[3]));
        v[0] = cadd(a0, b0);
        v[1] = cadd(a1, b1);
        v[2] = csub(a0, b0);
        v[3] = csub(a1, b1);
    }
};

// radix-2 split of an R-point DFT into two R/2-point DFTs (R = 8, 16, 32)
template <int R, bool INV> struct DFTsplit {
    __device__ __forceinline__ static void run(float2* v) {
        float2 e[R / 2], o[R / 2];
#pragma unroll
        for (int i = 0; i < R / 2; ++i) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
        DFT<R / 2, INV>::run(e);
        DFT<R / 2, INV>::run(o);
#pragma unroll
        for (int k = 0; k < R / 2; ++k) {
            float2 t;
            if (k == 0) t = o[0];
            else if (2 * k == R / 2) t = mul_mi<INV>(o[k]);
            else t = cmul(o[k], w64(k * (64 / R), INV));
            v[k] = cadd(e[k], t);
            v[k + R / 2] = csub(e[k], t);
        }
    }
};
template <bool INV> struct DFT<8, INV> : DFTsplit<8, INV> {};
template <bool INV> struct DFT<16, INV> : DFTsplit<16, INV> {};
template <bool INV> struct DFT<32, INV> : DFTsplit<32, INV> {};

template <bool INV> struct DFT<3, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        constexpr float s3 = 0.86602540378443865f;    // sin(2 pi / 3)
        const float2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        const float2 m = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
        // forward: -i sin(2pi/3) d ; inverse: +i sin(2pi/3) d
        const float2 r = INV ? make_float2(-s3 * d.y, s3 * d.x) : make_float2(s3 * d.y, -s3 * d.x);
        v[0] = cadd(v[0], s);
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    }
};

// ------------------------------------------------------------------ element points and buffer views
// An element as the engines hand it to functors: its coordinates and its index in the layout the
// current orientation streams (row-major [y][x] when consecutive lanes run along x, transposed [x][y]
// when they run along y).  The index is split as v + s, v varying per lane and s a constant of the
// unrolled element loop, so that buffer accesses take s as a scalar offset and cost no VALU.
struct Idx {
    unsigned v, s;
};
struct Pt {
    int y, x;
    Idx e;
};

// Raw buffer view of a global array: one scalar descriptor; an access at Idx (v, s) is a buffer
// instruction with VGPR offset v and SGPR offset s.  The hardware bounds-checks every access against
// the view's size: reads beyond it return 0 and writes beyond it are dropped.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
typedef unsigned u32x2 __attribute__((__vector_size__(8)));   // the type the 64-bit buffer builtins use
// (the pair is converted whole with bit_cast: element access on this vector type reads element 0 twice)
struct BufC {   // complex (float2) elements
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ BufC(const float2* p, unsigned n) : r(buffer_rsrc(p, n * 8u)) {}
    __device__ __forceinline__ float2 operator[](Idx e) const {
        return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, e.v * 8u, e.s * 8u, 0));
    }
    __device__ __forceinline__ void put(Idx e, float2 a) const {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, a), r, e.v * 8u, e.s * 8u, 0);
    }
};
struct BufF {   // real (float) elements
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ BufF(const float* p, unsigned n) : r(buffer_rsrc(p, n * 4u)) {}
    __device__ __forceinline__ float operator[](Idx e) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, e.v * 4u, e.s * 4u, 0));
    }
    __device__ __forceinline__ void put(Idx e, float a) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, a), r, e.v * 4u, e.s * 4u, 0);
    }
};

// ------------------------------------------------------------------ functors
struct NoOp {
    __device__ __forceinline__ float2 operator()(const Pt&, float2 v) const { return v; }
};

// An element functor split in two phases: load(p) fetches the element's global operands, apply(p, v,
// op) combines them with the element.  An engine that holds many elements per thread issues the loads
// of G elements back to back before the first apply, so their memory latencies overlap instead of
// adding up; used as a transform's post functor, the loads of its first PF elements may be issued
// before the transform's last pass, whose work then hides their latency.  Called as an ordinary
// functor, it runs both phases for one element.
template <int G_, class Ld, class Ap, int PF_ = 0, bool EAGER_ = false> struct Staged {
    static constexpr int G = G_;
    static constexpr int PF = PF_;   // elements whose loads an engine may issue ahead of the transform
    // issue the loads of every remaining element before the first apply, so that no load is issued
    // behind the stores an apply makes (a wait for it would also wait for those stores)
    static constexpr bool EAGER = EAGER_;
    Ld ld;
    Ap ap;
    __device__ __forceinline__ auto load(const Pt& p) const { return ld(p); }
    template <class V, class Op> __device__ __forceinline__ auto apply(const Pt& p, V&& v, const Op& o) const {
        return ap(p, v, o);
    }
    template <class V> __device__ __forceinline__ auto operator()(const Pt& p, V&& v) const { return ap(p, v, ld(p)); }
};
template <int G, int PF = 0, bool EAGER = false, class Ld, class Ap>
__device__ __forceinline__ Staged<G, Ld, Ap, PF, EAGER> staged(Ld ld, Ap ap) {
    return Staged<G, Ld, Ap, PF, EAGER>{ld, ap};
}
template <class F> struct IsStaged { static constexpr bool value = false; };
template <int G, class Ld, class Ap, int PF, bool EG> struct IsStaged<Staged<G, Ld, Ap, PF, EG>> {
    static constexpr bool value = true;
};
template <class F> struct PrefetchOf { static constexpr int value = 0; };
template <int G, class Ld, class Ap, int PF, bool EG> struct PrefetchOf<Staged<G, Ld, Ap, PF, EG>> {
    static constexpr int value = PF;
};

// ------------------------------------------------------------------ one Stockham stage
// ROWS: transform along x (lines are rows y); otherwise along y (lines are columns x).
// NS: span (product of the radices already done).  FIRST: apply `pre` on load (LOADG: pre supplies the
// value, LDS is not read).  LAST: apply `post` on store.
template <int N, int NT, int R, int NS, bool INV, bool ROWS, bool FIRST, bool LOADG, bool LAST, class Pre, class Post>
__device__ __forceinline__ void fft_stage(float2* __restrict__ s, const float2* __restrict__ tw, const Pre& pre,
                                          const Post& post) {
    constexpr int S = Pitch<N>::S;
    constexpr int NL = N / R;                 // butterflies per line
    constexpr int NB = N * NL;                // butterflies in the pass
    constexpr int CNT = (NB + NT - 1) / NT;   // butterflies per thread
    constexpr int TWS = N / (NS * R);         // twiddle table stride
    const int t = threadIdx.x;
    float2 v[CNT][R];
#pragma unroll
    for (int c = 0; c < CNT; ++c) {
        const int b = t + c * NT;
        if (NB % NT == 0 || b < NB) {
            const int line = b % N, j = b / N;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int e = j + r * NL;
                const int y = ROWS ? line : e, x = ROWS ? e : line;
                float2 val;
                const Pt pt{y, x, Idx{(unsigned)(ROWS ? x * N + y : y * N + x), 0u}};
                if constexpr (FIRST && LOADG) val = pre(pt, make_float2(0.f, 0.f));
                else if constexpr (FIRST) val = pre(pt, s[y * S + x]);
                else val = s[y * S + x];
                v[c][r] = val;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CNT; ++c) {
        const int b = t + c * NT;
        if (NB % NT == 0 || b < NB) {
            const int line = b % N, j = b / N;
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    float2 w = tw[k * r * TWS];
                    if (INV) w.y = -w.y;
                    v[c][r] = cmul(v[c][r], w);
                }
            }
            DFT<R, INV>::run(v[c]);
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int e = base + r * NS;
                const int y = ROWS ? line : e, x = ROWS ? e : line;
                float2 val = v[c][r];
                if constexpr (LAST) val = post(Pt{y, x, Idx{(unsigned)(ROWS ? x * N + y : y * N + x), 0u}}, val);
                s[y * S + x] = val;
            }
        }
    }
    __syncthreads();
}

template <int N, int NT, bool INV, bool ROWS, int ST, int NS, bool FIRSTPASS, bool LOADG, bool LASTPASS, class Pre,
          class Post>
__device__ __forceinline__ void fft_pass(float2* s, const float2* tw, const Pre& pre, const Post& post) {
    if constexpr (ST < Plan<N>::S) {
        constexpr int R = Plan<N>::R[ST];
        fft_stage<N, NT, R, NS, INV, ROWS, FIRSTPASS && ST == 0, LOADG, LASTPASS && ST == Plan<N>::S - 1>(s, tw, pre,
                                                                                                        post);
        fft_pass<N, NT, INV, ROWS, ST + 1, NS * R, FIRSTPASS, LOADG, LASTPASS>(s, tw, pre, post);
    }
}

// Forward 2D transform: columns (pre fused, lanes along x) then rows (post fused, lanes along y).
template <int N, int NT, bool LOADG = false, class Pre = NoOp, class Post = NoOp>
__device__ __forceinline__ void fft2d_fwd(float2* s, const float2* tw, const Pre& pre = Pre(), const Post& post = Post()) {
    fft_pass<N, NT, false, false, 0, 1, true, LOADG, false>(s, tw, pre, post);
    fft_pass<N, NT, false, true, 0, 1, false, false, true>(s, tw, pre, post);
}

// Inverse 2D transform (no 1/N^2): rows (pre fused, lanes along y) then columns (post fused, lanes along x).
template <int N, int NT, bool LOADG = false, class Pre = NoOp, class Post = NoOp>
__device__ __forceinline__ void fft2d_inv(float2* s, const float2* tw, const Pre& pre = Pre(), const Post& post = Post()) {
    fft_pass<N, NT, true, true, 0, 1, true, LOADG, false>(s, tw, pre, post);
    fft_pass<N, NT, true, false, 0, 1, false, false, true>(s, tw, pre, post);
}

// ------------------------------------------------------------------ batched 1D transforms of a tile of lines
// NLINES lines of N points in LDS, element e of line l at s[l * LS + e * ES]; consecutive lanes take
// consecutive lines at the same butterfly, so a pitch LS that is odd in float2 (ES = 1) or ES = NLINES
// with LS = 1 keeps the accesses free of bank conflicts.  Same radix plans and twiddle table as the
// plane transforms; an inverse transform leaves out the 1/N.
template <int N, int NT, int NLINES, int ES, int LS, int R, int NS, bool INV>
__device__ __forceinline__ void line_stage(float2* __restrict__ s, const float2* __restrict__ tw) {
    constexpr int NL = N / R;
    constexpr int NB = NLINES * NL;
    constexpr int CNT = (NB + NT - 1) / NT;
    constexpr int TWS = N / (NS * R);
    const int t = threadIdx.x;
    float2 v[CNT][R];
#pragma unroll
    for (int c = 0; c < CNT; ++c) {
        const int b = t + c * NT;
        if (NB % NT == 0 || b < NB) {
            const int line = b % NLINES, j = b / NLINES;   // lanes along lines: conflict-free with odd LS
#pragma unroll
            for (int r = 0; r < R; ++r) v[c][r] = s[line * LS + (j + r * NL) * ES];
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CNT; ++c) {
        const int b = t + c * NT;
        if (NB % NT == 0 || b < NB) {
            const int line = b % NLINES, j = b / NLINES;
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    float2 w = tw[k * r * TWS];
                    if (INV) w.y = -w.y;
                    v[c][r] = cmul(v[c][r], w);
                }
            }
            DFT<R, INV>::run(v[c]);
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) s[line * LS + (base + r * NS) * ES] = v[c][r];
        }
    }
    __syncthreads();
}

template <int N, int NT, int NLINES, int ES, int LS, int ST, int NS, bool INV>
__device__ __forceinline__ void line_fft(float2* s, const float2* tw) {
    if constexpr (ST < Plan<N>::S) {
        constexpr int R = Plan<N>::R[ST];
        line_stage<N, NT, NLINES, ES, LS, R, NS, INV>(s, tw);
        line_fft<N, NT, NLINES, ES, LS, ST + 1, NS * R, INV>(s, tw);
    }
}

// The LDS-stage transforms behind the engine interface shared with the register-resident engine:
// fwd / inv with (y, x, value) functors, each_R / each_C to visit the plane after a forward / inverse
// transform (lanes along y, resp. along x, so [x][y] resp. [y][x] global arrays stay coalesced).
template <int N, int NT_> struct LdsEngine {
    static constexpr int NT = NT_;
    static constexpr int S = Pitch<N>::S;
    static constexpr int LDS_ELEMS = N * S;
    static constexpr bool FUSED = false;   // no fused passes: the plane is on chip
    float2* s;
    const float2* tw;

    __device__ __forceinline__ void init(float2* lds, const float2* lds_tw) {
        s = lds;
        tw = lds_tw;
    }
    __device__ __forceinline__ void bind(float2*) {}   // the plane lives in LDS
    template <bool LOADG = false, class Pre = NoOp, class Post = NoOp>
    __device__ __forceinline__ void fwd(const Pre& pre = Pre(), const Post& post = Post()) {
        fft2d_fwd<N, NT, LOADG>(s, tw, pre, post);
    }
    template <bool LOADG = false, class Pre = NoOp, class Post = NoOp>
    __device__ __forceinline__ void inv(const Pre& pre = Pre(), const Post& post = Post()) {
        fft2d_inv<N, NT, LOADG>(s, tw, pre, post);
    }
    template <class F> __device__ __forceinline__ void each_R(const F& f) {
        for (int i = threadIdx.x; i < N * N; i += NT) {
            const int y = i % N, x = i / N;
            f(Pt{y, x, Idx{(unsigned)i, 0u}}, s[y * S + x]);
        }
        __syncthreads();
    }
    template <class F> __device__ __forceinline__ void each_C(const F& f) {
        for (int i = threadIdx.x; i < N * N; i += NT) {
            const int y = i / N, x = i % N;
            f(Pt{y, x, Idx{(unsigned)i, 0u}}, s[y * S + x]);
        }
        __syncthreads();
    }
};

}  // namespace mset
