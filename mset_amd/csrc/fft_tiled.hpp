// fft_tiled.hpp -- 2D FFT of a 256 x 256 complex plane that lives in global memory (CDNA4, gfx950).
//
// A 256 x 256 complex plane (512 KB) fits neither one CU's LDS nor its registers, so the plane stays in
// two per-workgroup global buffers and every 1D pass streams it through LDS in tiles of TL = 32 lines:
//   RS  row-major [y][x]     the state after an inverse transform (C orientation, lines are columns x)
//   XS  transposed [x][y]    the state after a forward transform  (R orientation, lines are rows y)
// A 256-point line transform is two radix-16 Stockham stages, one butterfly per thread per stage
// (32 lines x 16 butterflies = 512 threads).  The first stage reads global memory (applying `pre`), the
// second reads LDS and writes global memory (applying `post`), or writes LDS when the pass changes the
// layout, in which case the tile is copied out with the lanes along the other axis so that every global
// access stays coalesced.  The LDS tile is float2 s[pos][line] with pitch TL + 1 (odd): lanes along
// lines touch consecutive words, lanes along positions stride by 33 words, both free of bank conflicts.
//
//   fwd: column pass RS -> XS (pre fused on load, tile remapped on store), row pass XS -> XS (post fused)
//   inv: row pass XS -> RS (pre fused on load, remapped on store), column pass RS -> RS (post fused)
//
// Element points carry the index of the orientation's natural layout ([y][x] for columns, [x][y] for
// rows), as in the other engines, so the position routine runs unchanged.  each_R / each_C visit the
// plane read-only.
#pragma once
#include <hip/hip_runtime.h>
#include "fft_lds.hpp"

namespace mset {

template <int N, int NT_> struct TiledEngine {
    static constexpr int NT = NT_;
    static constexpr int TL = 32;                // lines per tile
    static constexpr int R = 16;                 // radix of both stages (N = 256)
    static constexpr int TP = TL + 1;            // LDS pitch
    static constexpr int LDS_ELEMS = N * TP;
    static_assert(N == 256 && NT == TL * (N / R), "one butterfly per thread per stage");
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

    // Transform of the 32 lines of one tile.  ld(line, pos) gives the input element, st(line, pos, v)
    // takes the output; in place in LDS when TO_LDS (st is then unused and the caller copies out).
    template <bool INV, bool TO_LDS, class Ld, class St> __device__ __forceinline__ void tile(const Ld& ld, const St& st) {
        const int t = threadIdx.x;
        const int line = t % TL, j = t / TL;
        float2 v[R];
        // stage 1 (span 1): butterfly j reads positions j + 16 r, writes 16 j + r
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = ld(line, j + R * r);
        DFT<R, INV>::run(v);
#pragma unroll
        for (int r = 0; r < R; ++r) s[(R * j + r) * TP + line] = v[r];
        __syncthreads();
        // stage 2 (span 16): butterfly j reads positions j + 16 r, twiddle W^(j r), writes j + 16 r
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = s[(j + R * r) * TP + line];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            const float2 w = tw[j * r];
            v[r] = INV ? cmulc(v[r], w) : cmul(v[r], w);
        }
        DFT<R, INV>::run(v);
        if constexpr (TO_LDS) {
#pragma unroll
            for (int r = 0; r < R; ++r) s[(j + R * r) * TP + line] = v[r];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) st(line, j + R * r, v[r]);
        }
        __syncthreads();
    }

    template <bool LOADG = false, class Pre = NoOp, class Post = NoOp>
    __device__ __forceinline__ void fwd(const Pre& pre = Pre(), const Post& post = Post()) {
        // column pass RS -> XS: lines are columns x = x0 + line, positions y
        for (int x0 = 0; x0 < N; x0 += TL) {
            tile<false, true>(
                [&](int line, int y) {
                    const int x = x0 + line;
                    const Pt p{y, x, Idx{(unsigned)(y * N + x), 0u}};
                    return pre(p, LOADG ? make_float2(0.f, 0.f) : RS[y * N + x]);
                },
                NoOp());
            for (int e = threadIdx.x; e < TL * N; e += NT) {         // lanes along y: XS[x][y]
                const int y = e % N, l = e / N;
                XS[(x0 + l) * N + y] = s[y * TP + l];
            }
            __syncthreads();
        }
        // row pass XS -> XS: lines are rows y = y0 + line, positions x
        for (int y0 = 0; y0 < N; y0 += TL) {
            tile<false, false>([&](int line, int x) { return XS[x * N + y0 + line]; },
                               [&](int line, int x, float2 v) {
                                   const int y = y0 + line;
                                   XS[x * N + y] = post(Pt{y, x, Idx{(unsigned)(x * N + y), 0u}}, v);
                               });
        }
    }

    template <bool LOADG = false, class Pre = NoOp, class Post = NoOp>
    __device__ __forceinline__ void inv(const Pre& pre = Pre(), const Post& post = Post()) {
        // row pass XS -> RS
        for (int y0 = 0; y0 < N; y0 += TL) {
            tile<true, true>(
                [&](int line, int x) {
                    const int y = y0 + line;
                    const Pt p{y, x, Idx{(unsigned)(x * N + y), 0u}};
                    return pre(p, LOADG ? make_float2(0.f, 0.f) : XS[x * N + y]);
                },
                NoOp());
            for (int e = threadIdx.x; e < TL * N; e += NT) {         // lanes along x: RS[y][x]
                const int x = e % N, l = e / N;
                RS[(y0 + l) * N + x] = s[x * TP + l];
            }
            __syncthreads();
        }
        // column pass RS -> RS
        for (int x0 = 0; x0 < N; x0 += TL) {
            tile<true, false>([&](int line, int y) { return RS[y * N + x0 + line]; },
                              [&](int line, int y, float2 v) {
                                  const int x = x0 + line;
                                  RS[y * N + x] = post(Pt{y, x, Idx{(unsigned)(y * N + x), 0u}}, v);
                              });
        }
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
