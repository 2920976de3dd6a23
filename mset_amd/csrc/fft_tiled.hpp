// fft_tiled.hpp -- 2D FFT of a 256 x 256 complex plane that lives in global memory (CDNA4, gfx950).
//
// A 256 x 256 complex plane (512 KB) fits neither one CU's LDS nor its registers, so the plane stays in
// two per-workgroup global buffers and every 1D pass streams it through LDS in tiles of TL = 32 lines:
//   RS  row-major [y][x]     the state after an inverse transform (C orientation, lines are columns x)
//   XS  transposed [x][y]    the state after a forward transform  (R orientation, lines are rows y)
// A 256-point line transform is two radix-16 Stockham stages, one butterfly per thread per stage
// (32 lines x 16 butterflies = 512 threads), exchanging through LDS.  A pass loads a tile from global
// memory (applying `pre`), runs up to two line transforms with an element functor between them, and
// stores the tile back in place or, when the pass changes the layout, through LDS with the lanes along
// the other axis so that every global access stays coalesced.  The LDS tile is float2 s[pos][line] with
// pitch TL + 1 (odd): lanes along lines touch consecutive words, lanes along positions stride by 33
// words, both free of bank conflicts.
//
//   fwd: column pass RS -> XS (pre fused on load), row pass XS -> XS (post fused)
//   inv: row pass XS -> RS (pre fused on load), column pass RS -> RS (post fused)
// A sweep of propagation steps fuses the passes that meet between transforms: [rowF, x p, rowI] and
// [colI, real-space product, colF] each stream the plane once, so a step costs two reads and two
// writes of the plane instead of four.
//
// Element points carry the index of the orientation's natural layout ([y][x] for columns, [x][y] for
// rows), as in the other engines, so the position routine runs unchanged.  each_R / each_C visit the
// plane read-only.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include "fft_lds.hpp"

namespace mset {

template <int N, int NT_> struct TiledEngine {
    static constexpr int NT = NT_;
    static constexpr int TL = 32;                // lines per tile
    static constexpr int R = 16;                 // radix of both stages (N = 256)
    static constexpr int TP = TL + 1;            // LDS pitch
    static constexpr int LDS_ELEMS = N * TP;
    static constexpr bool FUSED = true;          // provides the fused passes below
    static_assert(N == 256 && NT == TL * (N / R), "one butterfly per thread per stage");
    enum { NONE = 0, FWD = 1, INV = 2 };
    float2* s;
    const float2* tw;
    float2* RS;   // [y][x]
    float2* XS;   // [x][y]

    __device__ __forceinline__ void init(float2* lds, const float2* lds_tw) {
        s = lds;
        tw = lds_tw;
    }
    // the workgroup's two plane buffers (2 N^2 elements)
    __device__ __forceinline__ void bind(float2* planes) {
        RS = planes;
        XS = planes + N * N;
    }

    // 256-point transform of the thread's elements j + 16 r (r < 16) of line `line`, in and out; LEAD: the
    // LDS may still be read by other threads (a transform earlier in the same tile) and needs a barrier
    template <int D, bool LEAD> __device__ __forceinline__ void dft(float2 (&v)[R], int line, int j) {
        // stage 1 (span 1): butterfly j over positions j + 16 r, written to 16 j + r
        DFT<R, D == INV>::run(v);
        if constexpr (LEAD) __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) s[(R * j + r) * TP + line] = v[r];
        __syncthreads();
        // stage 2 (span 16): butterfly j reads positions j + 16 r, twiddle W^(j r), writes j + 16 r
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = s[(j + R * r) * TP + line];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            const float2 w = tw[j * r];
            v[r] = D == INV ? cmulc(v[r], w) : cmul(v[r], w);
        }
        DFT<R, D == INV>::run(v);
    }

    // One pass over the plane in tiles of 32 lines: load (through `pre`; LOADG: pre supplies the value),
    // transform D1, `mid` on every element, transform D2, store.  COL: lines are columns x, positions y,
    // the input is RS; otherwise lines are rows y, positions x, the input is XS.  SWAP: the output goes to
    // the other buffer (through LDS, lanes along its fast axis); otherwise back in place.  Points carry
    // the index of the orientation's natural layout ([y][x] for columns, [x][y] for rows).
    template <bool COL, int D1, int D2, bool LOADG, bool SWAP, class Pre, class Mid>
    __device__ __forceinline__ void pass(const Pre& pre, const Mid& mid) {
        const int t = threadIdx.x;
        const int line = t % TL, j = t / TL;
        float2* in = COL ? RS : XS;
        float2* out = SWAP ? (COL ? XS : RS) : in;
        for (int l0 = 0; l0 < N; l0 += TL) {
            const int ln = l0 + line;
            auto point = [&](int pos) {
                return COL ? Pt{pos, ln, Idx{(unsigned)(pos * N + ln), 0u}} : Pt{ln, pos, Idx{(unsigned)(pos * N + ln), 0u}};
            };
            float2 v[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int pos = j + R * r;
                if constexpr (std::is_same<Pre, NoOp>::value && !LOADG) v[r] = in[pos * N + ln];
                else v[r] = pre(point(pos), LOADG ? make_float2(0.f, 0.f) : in[pos * N + ln]);
            }
            if constexpr (IsStaged<Mid>::value) {
                // all 16 elements' operands are loaded before the first apply, so their latencies overlap
                // (loading them ahead of D1 instead measured slower: the registers spill)
                if constexpr (D1 != NONE) dft<D1, false>(v, line, j);
                decltype(mid.load(Pt{})) op[R];
#pragma unroll
                for (int r = 0; r < R; ++r) op[r] = mid.load(point(j + R * r));
#pragma unroll
                for (int r = 0; r < R; ++r) v[r] = mid.apply(point(j + R * r), v[r], op[r]);
            } else {
                if constexpr (D1 != NONE) dft<D1, false>(v, line, j);
                if constexpr (!std::is_same<Mid, NoOp>::value) {
#pragma unroll
                    for (int r = 0; r < R; ++r) v[r] = mid(point(j + R * r), v[r]);
                }
            }
            if constexpr (D2 != NONE) dft<D2, D1 != NONE>(v, line, j);
            if constexpr (SWAP) {
                // the tile to LDS as s[pos][line], then out along the other axis
                // each thread rewrites the LDS words it read last (or, with no transform, words free
                // since the previous tile's closing barrier): no barrier needed before
#pragma unroll
                for (int r = 0; r < R; ++r) s[(j + R * r) * TP + line] = v[r];
                __syncthreads();
                for (int e = t; e < TL * N; e += NT) {
                    const int pos = e % N, l = e / N;
                    out[(l0 + l) * N + pos] = s[pos * TP + l];
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) out[(j + R * r) * N + ln] = v[r];
            }
            __syncthreads();
        }
    }

    // The four 1D passes of a 2D transform and the plane's states between them:
    //   real (RS) --colF--> half-forward (XS) --rowF--> Fourier (XS) --rowI--> half-inverse (RS) --colI--> real
    template <bool LOADG = false, class Pre = NoOp, class Post = NoOp>
    __device__ __forceinline__ void fwd(const Pre& pre = Pre(), const Post& post = Post()) {
        pass<true, FWD, NONE, LOADG, true>(pre, NoOp());
        pass<false, FWD, NONE, false, false>(NoOp(), post);
    }
    template <bool LOADG = false, class Pre = NoOp, class Post = NoOp>
    __device__ __forceinline__ void inv(const Pre& pre = Pre(), const Post& post = Post()) {
        pass<false, INV, NONE, LOADG, true>(pre, NoOp());
        pass<true, INV, NONE, false, false>(NoOp(), post);
    }

    // Fused passes for a sweep of propagation steps: each streams the plane once instead of twice.
    // row_step: half-forward -> [rowF, mid, rowI] -> half-inverse (the Fourier-domain product of a step)
    template <class Mid> __device__ __forceinline__ void row_step(const Mid& mid) {
        pass<false, FWD, INV, false, true>(NoOp(), mid);
    }
    // col_step: half-inverse -> [colI, post, colF] -> half-forward (the real-space product between steps)
    template <class Post> __device__ __forceinline__ void col_step(const Post& post) {
        pass<true, INV, FWD, false, true>(NoOp(), post);
    }
    // row_last: half-forward -> [rowF, post] -> Fourier (the end of a forward transform)
    template <class Post> __device__ __forceinline__ void row_last(const Post& post) {
        pass<false, FWD, NONE, false, false>(NoOp(), post);
    }
    // row_first: Fourier -> [pre, rowI] -> half-inverse (the start of an inverse transform)
    template <bool LOADG = false, class Pre = NoOp> __device__ __forceinline__ void row_first(const Pre& pre) {
        pass<false, INV, NONE, LOADG, true>(pre, NoOp());
    }

    // col_first: real -> [pre, colF] -> half-forward (the start of a forward transform)
    template <bool LOADG = false, class Pre = NoOp> __device__ __forceinline__ void col_first(const Pre& pre) {
        pass<true, FWD, NONE, LOADG, true>(pre, NoOp());
    }
    // col_last: half-inverse -> [colI, post] -> real (the end of an inverse transform)
    template <class Post> __device__ __forceinline__ void col_last(const Post& post) {
        pass<true, INV, NONE, false, false>(NoOp(), post);
    }
    // Rebinds the row-major plane buffer, returning the previous one: the state held there stays put
    // while the engine works on another plane (the recompute variant parks the adjoint wave this way).
    __device__ __forceinline__ float2* swap_rs(float2* other) {
        float2* old = RS;
        RS = other;
        return old;
    }

    template <class F> __device__ __forceinline__ void each_R(const F& f) {
        for (int e = threadIdx.x; e < N * N; e += NT) {
            const int y = e % N, x = e / N;
            float2 v = XS[e];
            f(Pt{y, x, Idx{(unsigned)e, 0u}}, v);
        }
        __syncthreads();
    }
    template <class F> __device__ __forceinline__ void each_C(const F& f) {
        for (int e = threadIdx.x; e < N * N; e += NT) {
            const int y = e / N, x = e % N;
            float2 v = RS[e];
            f(Pt{y, x, Idx{(unsigned)e, 0u}}, v);
        }
        __syncthreads();
    }
};

}  // namespace mset
