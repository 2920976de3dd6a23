// fft_tiled.hpp -- 2D FFT of a 256 x 256 complex plane that lives in global memory (CDNA4, gfx950).
//
// A 256 x 256 complex plane (512 KB) fits neither one CU's LDS nor its registers, so the plane stays in
// two per-workgroup global buffers and every 1D pass streams it through LDS in tiles of TL lines (TL = NT / 16:
// 32 with 512 threads, 16 with 256, 64 with 1024):
//   RS  row-major [y][x]     the state after an inverse transform (C orientation, lines are columns x)
//   XS  transposed [x][y]    the state after a forward transform  (R orientation, lines are rows y)
// A 256-point line transform is two radix-16 Stockham stages, one butterfly per thread per stage
// (TL lines x 16 butterflies = NT threads), exchanging through LDS.  A pass loads a tile from global
// memory (applying `pre`), runs up to two line transforms with an element functor between them, and
// stores the tile back in place or, when the pass changes the layout, through LDS with the lanes along
// the other axis so that every global access stays coalesced.  The LDS tile is float2 s[pos][line] with
// pitch TL + 1 (odd): lanes along lines touch consecutive words, lanes along positions stride by TL + 1
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

// Tile steps of a chain pass (TiledEngine::chain): they exchange the whole tile with a global array kept
// in the OTHER orientation's layout, through the LDS tile, so that access is coalesced as well.
template <class F> struct IsTileStep { static constexpr bool value = false; };

// stores the tile's current values (unchanged) into dst; no-op when dst is null
struct TileStore {
    float2* dst;
    template <bool COL, class E, int R> __device__ __forceinline__ void run(E& e, float2 (&v)[R], int l0, int line,
                                                                            int j) const {
        if (dst) e.tile_to_other(v, dst, l0, line, j);
    }
};
template <> struct IsTileStep<TileStore> { static constexpr bool value = true; };

// loads one operand per element from src and combines it: v = ap(point, v, operand)
template <class T, class Ap> struct TileLoad {
    const T* src;
    Ap ap;
    template <bool COL, class E, int R> __device__ __forceinline__ void run(E& e, float2 (&v)[R], int l0, int line,
                                                                            int j) const {
        T o[R];
        e.other_to_tile(src, o, l0, line, j);
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = ap(e.template pt<COL>(l0, line, j + R * r), v[r], o[r]);
    }
};
template <class T, class Ap> struct IsTileStep<TileLoad<T, Ap>> { static constexpr bool value = true; };
template <class T, class Ap> __device__ __forceinline__ TileLoad<T, Ap> tile_load(const T* src, Ap ap) {
    return TileLoad<T, Ap>{src, ap};
}

// several steps applied in order as one step of a chain pass
template <class A, class B, class C = NoOp> struct Seq {
    A a;
    B b;
    C c;
};
template <class F> struct IsSeq { static constexpr bool value = false; };
template <class A, class B, class C> struct IsSeq<Seq<A, B, C>> { static constexpr bool value = true; };
template <class A, class B> __device__ __forceinline__ Seq<A, B> seq(A a, B b) { return Seq<A, B>{a, b, NoOp()}; }
template <class A, class B, class C> __device__ __forceinline__ Seq<A, B, C> seq(A a, B b, C c) {
    return Seq<A, B, C>{a, b, c};
}
template <class F> struct HasTileStep { static constexpr bool value = IsTileStep<F>::value; };
template <class A, class B, class C> struct HasTileStep<Seq<A, B, C>> {
    static constexpr bool value = HasTileStep<A>::value || HasTileStep<B>::value || HasTileStep<C>::value;
};

template <int N, int NT_> struct TiledEngine {
    static constexpr int NT = NT_;
    static constexpr int R = 16;                 // radix of both stages (N = 256)
    static constexpr int TL = NT / (N / R);      // lines per tile
    static constexpr int TP = TL + 1;            // LDS pitch
    static constexpr int LDS_ELEMS = N * TP;
    static constexpr bool FUSED = true;          // provides the fused passes below
    static_assert(N == 256 && NT == TL * (N / R) && (TL == 16 || TL == 32 || TL == 64), "one butterfly per thread per stage");
    enum { NONE = 0, FWD = 1, INV = 2 };
    float2* s;
    const float2* tw;
    float2* RS;   // [y][x]
    float2* XS;   // [x][y]
    float2* PB;   // the workgroup's plane buffers (bind)

    __device__ __forceinline__ void init(float2* lds, const float2* lds_tw) {
        s = lds;
        tw = lds_tw;
    }
    // the workgroup's two plane buffers (2 N^2 elements)
    __device__ __forceinline__ void bind(float2* planes) {
        PB = planes;
        RS = planes;
        XS = planes + N * N;
    }
    // the next chain pass of orientation COL reads its tile from `in` and writes it to `out` (each in the
    // layout the plane buffer it replaces has; null: that plane buffer); unbind() goes back to the buffers
    template <bool COL> __device__ __forceinline__ void redirect(const float2* in, float2* out) {
        unbind();
        float2*& src = COL ? RS : XS;
        float2*& dst = COL ? XS : RS;
        if (in) src = const_cast<float2*>(in);
        if (out) dst = out;
    }
    __device__ __forceinline__ void unbind() {
        RS = PB;
        XS = PB + N * N;
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

    // One pass over the plane in tiles of TL lines: load (through `pre`; LOADG: pre supplies the value),
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

    // ---------------------------------------------------------------- chain passes
    // A chain pass runs up to four line transforms D1..D4 on each tile, with a step between consecutive
    // transforms (and after the last): the tile is loaded once (through `pre`), stored once (SWAP: to the
    // other buffer through LDS; NONE: not at all), and everything between stays in registers.  Steps are
    // NoOp, plain element functors, staged functors (their global operands loaded for the whole tile
    // first) or tile steps (TileLoad / TileStore below), which exchange the tile with a global array in
    // the OTHER orientation's layout through the LDS tile, so that access stays coalesced too.
    enum { ST_NONE = 0, ST_SWAP = 1 };

    template <bool COL> __device__ __forceinline__ Pt pt(int l0, int line, int pos) const {
        const int ln = l0 + line;
        return COL ? Pt{pos, ln, Idx{(unsigned)(pos * N + ln), 0u}} : Pt{ln, pos, Idx{(unsigned)(pos * N + ln), 0u}};
    }

    // one step applied to the thread's 16 elements of the tile
    template <bool COL, class M> __device__ __forceinline__ void step(const M& m, float2 (&v)[R], int l0, int line, int j) {
        if constexpr (std::is_same<M, NoOp>::value) {
            return;
        } else if constexpr (IsSeq<M>::value) {
            step<COL>(m.a, v, l0, line, j);
            step<COL>(m.b, v, l0, line, j);
            step<COL>(m.c, v, l0, line, j);
        } else if constexpr (IsTileStep<M>::value) {
            m.template run<COL>(*this, v, l0, line, j);
        } else if constexpr (IsStaged<M>::value) {
            // operands in groups of G elements: each group's loads are issued back to back before its applies
            constexpr int G = M::G < R ? M::G : R;
            static_assert(R % G == 0, "staging group must divide the thread's elements");
#pragma unroll
            for (int r0 = 0; r0 < R; r0 += G) {
                decltype(m.load(Pt{})) op[G];
#pragma unroll
                for (int r = 0; r < G; ++r) op[r] = m.load(pt<COL>(l0, line, j + R * (r0 + r)));
#pragma unroll
                for (int r = 0; r < G; ++r) v[r0 + r] = m.apply(pt<COL>(l0, line, j + R * (r0 + r)), v[r0 + r], op[r]);
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = m(pt<COL>(l0, line, j + R * r), v[r]);
        }
    }

    template <bool COL, int ST, int D1, int D2 = NONE, int D3 = NONE, int D4 = NONE, class P0 = NoOp, class M1 = NoOp,
              class M2 = NoOp, class M3 = NoOp, class M4 = NoOp>
    __device__ __forceinline__ void chain(const P0& pre, const M1& m1 = M1(), const M2& m2 = M2(), const M3& m3 = M3(),
                                          const M4& m4 = M4()) {
        const int t = threadIdx.x;
        const int line = t % TL, j = t / TL;
        const float2* in = COL ? RS : XS;
        float2* out = COL ? XS : RS;
        constexpr bool PRE_TILE = IsTileStep<P0>::value;   // a tile load that supplies the values
        for (int l0 = 0; l0 < N; l0 += TL) {
            float2 v[R];
            if constexpr (PRE_TILE) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[r] = make_float2(0.f, 0.f);
                pre.template run<COL>(*this, v, l0, line, j);
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) v[r] = in[(j + R * r) * N + l0 + line];
                step<COL>(pre, v, l0, line, j);
            }
            if constexpr (D1 != NONE) dft<D1, PRE_TILE>(v, line, j);
            step<COL>(m1, v, l0, line, j);
            if constexpr (D2 != NONE) dft<D2, true>(v, line, j);
            step<COL>(m2, v, l0, line, j);
            if constexpr (D3 != NONE) dft<D3, true>(v, line, j);
            step<COL>(m3, v, l0, line, j);
            if constexpr (D4 != NONE) dft<D4, true>(v, line, j);
            step<COL>(m4, v, l0, line, j);
            if constexpr (ST == ST_SWAP) {
                // after a transform each thread rewrites the LDS words it read last; after a tile step the
                // words differ, so the tile step's readers must be done first
                constexpr bool T1 = HasTileStep<M1>::value, T2 = HasTileStep<M2>::value;
                constexpr bool T3 = HasTileStep<M3>::value, T4 = HasTileStep<M4>::value;
                constexpr bool TILE_LAST = (D4 != NONE) ? T4
                                           : (D3 != NONE) ? (T3 || T4)
                                           : (D2 != NONE) ? (T2 || T3 || T4)
                                           : (D1 != NONE) ? (T1 || T2 || T3 || T4)
                                           : true;
                if constexpr (TILE_LAST) __syncthreads();
#pragma unroll
                for (int r = 0; r < R; ++r) s[(j + R * r) * TP + line] = v[r];
                __syncthreads();
                for (int e = t; e < TL * N; e += NT) {
                    const int pos = e % N, l = e / N;
                    out[(l0 + l) * N + pos] = s[pos * TP + l];
                }
            }
            __syncthreads();
        }
    }

    // LDS round trip of a tile between the thread layout (16 elements of one line) and a global array in
    // the other orientation's layout (lines slow, positions fast): element (line l0 + l, position pos) of
    // the array is dst[(l0 + l) N + pos].  Callers ensure the LDS tile is free (a barrier, or a transform's
    // LEAD barrier after).
    __device__ __forceinline__ void tile_to_other(const float2 (&v)[R], float2* dst, int l0, int line, int j) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) s[(j + R * r) * TP + line] = v[r];
        __syncthreads();
        for (int e = threadIdx.x; e < TL * N; e += NT) {
            const int pos = e % N, l = e / N;
            dst[(l0 + l) * N + pos] = s[pos * TP + l];
        }
    }
    template <class T> __device__ __forceinline__ void other_to_tile(const T* src, T (&o)[R], int l0, int line, int j) {
        __syncthreads();
        T* st = reinterpret_cast<T*>(s);   // T is float or float2: at most one float2 per LDS slot
        for (int e = threadIdx.x; e < TL * N; e += NT) {
            const int pos = e % N, l = e / N;
            st[(pos * TP + l) * (sizeof(float2) / sizeof(T))] = src[(l0 + l) * N + pos];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) o[r] = st[((j + R * r) * TP + line) * (sizeof(float2) / sizeof(T))];
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
