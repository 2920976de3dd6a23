// fft_reg.hpp -- register-resident 2D FFT of a 128 x 128 complex plane for CDNA4 (gfx950).
//
// The workgroup has 512 threads (8 waves).  In a wave, lane = 16 h + c: the four lanes of one column c
// (one per 16-lane row h) own one line of the plane (a column in the "C" orientation, a row in the "R"
// orientation), 32 elements each, in registers; a wave owns 16 lines, line = 16 wave + c.
//
// A 128-point transform is split 128 = 4 x 32.  Decimation in time (forward):
//   row h holds x[h + 4 m] (register m)                                   layout D
//   -> 32-point DFT over m in registers, twiddle W_128^(h k)
//   -> per group of four registers k = 4 g .. 4 g + 3, a 4 x 4 transpose across the rows with two
//      v_permlane32_swap and two v_permlane16_swap per component (row p ends up with the four rows'
//      values of k = 4 g + p) and a 4-point DFT in registers:
//   row h holds X[32 j + 4 g + h] in register 4 g + j                     layout X
// Decimation in frequency (inverse) runs the same steps backwards and maps layout X to layout D, so a
// forward transform's output is exactly what the inverse consumes and vice versa: consecutive
// transforms never reshuffle registers.  No LDS, no barrier and no lane-dependent select is involved
// in a 1D transform; the twiddles W_128^(h k) come from an LDS table (the 16 lanes of a row read the
// same entry: a broadcast).
//
// A 2D transform is a column pass, one transpose through LDS (C <-> R) and a row pass.  The transpose
// buffer has an odd row pitch (129 float2): the 16 lanes of a row access either 16 consecutive
// elements of one row or the same element of 16 consecutive rows, both free of bank conflicts.
//
// Interface (shared with LdsEngine in fft_lds.hpp): fwd(pre, post) goes C -> R and inv(pre, post)
// goes R -> C; pre / post see (point, value) for every element the thread owns; each_C / each_R visit
// the owned elements after an inverse / forward transform.  In C, the 16 lanes of a row hold 16
// consecutive x (coalesced for row-major [y][x] arrays); in R they hold 16 consecutive y (coalesced
// for arrays stored [x][y]).
#pragma once
#include <hip/hip_runtime.h>
#include <utility>
#include "fft_lds.hpp"

namespace mset {

// exchange across 16-lane rows: rows 2, 3 of a <-> rows 0, 1 of b (v_permlane32_swap), or odd rows of
// a <-> even rows of b (v_permlane16_swap)
__device__ __forceinline__ void swap32(float& a, float& b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, a), __builtin_bit_cast(unsigned, b),
                                                    false, false);
    a = __builtin_bit_cast(float, (unsigned)r[0]);
    b = __builtin_bit_cast(float, (unsigned)r[1]);
}
__device__ __forceinline__ void swap16(float& a, float& b) {
    const auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, a), __builtin_bit_cast(unsigned, b),
                                                    false, false);
    a = __builtin_bit_cast(float, (unsigned)r[0]);
    b = __builtin_bit_cast(float, (unsigned)r[1]);
}

// 4 x 4 transpose across the rows of a wave: before, row p register q holds M[p][q]; after, row p
// register q holds M[q][p]
__device__ __forceinline__ void transpose4(float2* v) {
    swap32(v[0].x, v[2].x);
    swap32(v[0].y, v[2].y);
    swap32(v[1].x, v[3].x);
    swap32(v[1].y, v[3].y);
    swap16(v[0].x, v[1].x);
    swap16(v[0].y, v[1].y);
    swap16(v[2].x, v[3].x);
    swap16(v[2].y, v[3].y);
}

template <int N, bool HALF = false> struct RegEngine;

// HALF: the transposes go through a buffer of half the size, real parts then imaginary parts (66 KB
// instead of 132 KB), so that two workgroups fit in a CU's LDS
template <bool HALF> struct RegEngine<128, HALF> {
    static constexpr int N = 128, E = 32, NT = 512, SP = 129;
    static constexpr int LDS_ELEMS = HALF ? (N * SP + 1) / 2 : N * SP;
    // the half-LDS engine runs two workgroups per CU at 128 registers: staged functors load their operands
    // at most GMAX elements ahead and nothing ahead of the transform
    static constexpr int GMAX = HALF ? 4 : E;
    static constexpr bool AHEAD = !HALF;
    static constexpr bool FUSED = false;   // no fused passes: the plane is in registers
    float2* s;
    const float2* tw;  // W_128^k table in LDS
    float2 v[E];
    int line, h;

    __device__ __forceinline__ void init(float2* lds, const float2* lds_tw) {
        s = lds;
        tw = lds_tw;
        const int t = threadIdx.x;
        line = ((t >> 6) << 4) | (t & 15);
        h = (t >> 4) & 3;
    }
    __device__ __forceinline__ void bind(float2*) {}   // the plane lives in registers

    static __device__ __forceinline__ int A(int y, int x) { return y * SP + x; }
    // element index along the line of register i in layout X (layout D: h + 4 i)
    __device__ __forceinline__ int xidx(int i) const { return 32 * (i & 3) + 4 * (i >> 2) + h; }

    // The element functors load global operands; unrolled over 32 elements the scheduler would hoist
    // every load ahead of its use and run out of registers.  A scheduling fence after every G
    // elements bounds the work in flight to G elements' worth.
    template <int G = 4> static __device__ __forceinline__ void group_fence(int i) {
        if ((i & (G - 1)) == G - 1) __builtin_amdgcn_sched_barrier(0);
    }

    // Per-element addresses are cheap affine functions of (line, h).  Left visible, the compiler hoists
    // all of them out of the plane and position loops and keeps hundreds of registers live; routing the
    // coordinates through an empty asm at every transform makes them opaque, so the addresses are
    // recomputed where they are used.
    __device__ __forceinline__ void opaque() { asm volatile("" : "+v"(line), "+v"(h)); }

    // layout D -> layout X
    template <bool INV> __device__ __forceinline__ void pass_dit() {
        DFT<E, INV>::run(v);
#pragma unroll
        for (int k = 1; k < E; ++k) {
            const float2 w = tw[h * k];
            v[k] = INV ? cmulc(v[k], w) : cmul(v[k], w);
            group_fence<8>(k);
        }
#pragma unroll
        for (int g = 0; g < E / 4; ++g) {
            transpose4(v + 4 * g);
            DFT<4, INV>::run(v + 4 * g);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // layout X -> layout D
    template <bool INV> __device__ __forceinline__ void pass_dif() {
#pragma unroll
        for (int g = 0; g < E / 4; ++g) {
            DFT<4, INV>::run(v + 4 * g);
            transpose4(v + 4 * g);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int m = 1; m < E; ++m) {
            const float2 w = tw[h * m];
            v[m] = INV ? cmulc(v[m], w) : cmul(v[m], w);
            group_fence<8>(m);
        }
        DFT<E, INV>::run(v);
    }

    // C orientation, layout X (y = xidx, x = line) -> R orientation, layout D (y = line, x = h + 4 m)
    __device__ __forceinline__ void transpose_c_to_r() {
        if constexpr (HALF) {
            // a component's old values are dead once written, so the new ones take their registers
            float* f = reinterpret_cast<float*>(s);
            __syncthreads();
#pragma unroll
            for (int i = 0; i < E; ++i) f[A(xidx(i), line)] = v[i].x;
            __syncthreads();
#pragma unroll
            for (int m = 0; m < E; ++m) v[m].x = f[A(line, h + 4 * m)];
            __syncthreads();
#pragma unroll
            for (int i = 0; i < E; ++i) f[A(xidx(i), line)] = v[i].y;
            __syncthreads();
#pragma unroll
            for (int m = 0; m < E; ++m) v[m].y = f[A(line, h + 4 * m)];
        } else {
            __syncthreads();
#pragma unroll
            for (int i = 0; i < E; ++i) s[A(xidx(i), line)] = v[i];
            __syncthreads();
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = s[A(line, h + 4 * m)];
        }
    }
    // R orientation, layout D (y = line, x = h + 4 m) -> C orientation, layout X (y = xidx, x = line)
    __device__ __forceinline__ void transpose_r_to_c() {
        if constexpr (HALF) {
            float* f = reinterpret_cast<float*>(s);
            __syncthreads();
#pragma unroll
            for (int m = 0; m < E; ++m) f[A(line, h + 4 * m)] = v[m].x;
            __syncthreads();
#pragma unroll
            for (int i = 0; i < E; ++i) v[i].x = f[A(xidx(i), line)];
            __syncthreads();
#pragma unroll
            for (int m = 0; m < E; ++m) f[A(line, h + 4 * m)] = v[m].y;
            __syncthreads();
#pragma unroll
            for (int i = 0; i < E; ++i) v[i].y = f[A(xidx(i), line)];
        } else {
            __syncthreads();
#pragma unroll
            for (int m = 0; m < E; ++m) s[A(line, h + 4 * m)] = v[m];
            __syncthreads();
#pragma unroll
            for (int i = 0; i < E; ++i) v[i] = s[A(xidx(i), line)];
        }
    }

    // Visit the 32 owned elements.  COL: C orientation, layout D, element i is (h + 4 i, line), index
    // (h N + line) + 4 N i in a row-major [y][x] array; otherwise R orientation, layout X, element i is
    // (line, xidx(i)), index (h N + line) + (32 (i & 3) + 4 (i >> 2)) N in a transposed [x][y] array.
    // MAP: the functor returns the new value (LOADG: it is handed zero instead of the current one);
    // otherwise it sees the element by reference.
    template <bool COL> __device__ __forceinline__ Pt point(int i, unsigned vb) const {
        if (COL) return Pt{h + 4 * i, line, Idx{vb, (unsigned)(4 * N * i)}};
        return Pt{line, xidx(i), Idx{vb, (unsigned)((32 * (i & 3) + 4 * (i >> 2)) * N)}};
    }
    __device__ __forceinline__ unsigned vbase() const { return (unsigned)(h * N + line); }

    // operands of the first PF elements of a staged functor, loaded ahead (see Staged)
    template <class F, int PF = PrefetchOf<F>::value> struct Ahead {
        decltype(std::declval<const F&>().load(Pt{})) op[PF > 0 ? PF : 1];
    };
    template <bool COL, class F, class AH> __device__ __forceinline__ void load_ahead(const F& f, AH& ah) {
        constexpr int PF = PrefetchOf<F>::value;
        if constexpr (PF > 0) {
            const unsigned vb = vbase();
#pragma unroll
            for (int i = 0; i < PF; ++i) ah.op[i] = f.load(point<COL>(i, vb));
            __builtin_amdgcn_sched_barrier(0);
        }
    }

    template <bool COL, bool MAP, bool LOADG, class F> __device__ __forceinline__ void run_elems(const F& f) {
        if constexpr (IsStaged<F>::value) {
            Ahead<F, 0> none;
            run_staged<COL, MAP, LOADG, 0>(f, none);
        } else {
            const unsigned vb = vbase();
#pragma unroll
            for (int i = 0; i < E; ++i) {
                if constexpr (MAP) v[i] = f(point<COL>(i, vb), LOADG ? make_float2(0.f, 0.f) : v[i]);
                else f(point<COL>(i, vb), v[i]);
                group_fence(i);
            }
        }
    }
    // a staged functor over the 32 elements; the operands of the first PF are already in `ah`
    template <bool COL, bool MAP, bool LOADG, int PF, class F, class AH>
    __device__ __forceinline__ void run_staged(const F& f, const AH& ah) {
        constexpr int G = F::G < GMAX ? F::G : GMAX;
        static_assert(E % G == 0 && PF % G == 0 && PF <= E, "group size must divide 32 and the prefetch");
        const unsigned vb = vbase();
        if constexpr (F::EAGER && AHEAD && PF < E) {
            decltype(f.load(Pt{})) rest[E - PF];
#pragma unroll
            for (int i = PF; i < E; ++i) rest[i - PF] = f.load(point<COL>(i, vb));
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int g0 = 0; g0 < E; g0 += G) {
#pragma unroll
                for (int i = 0; i < G; ++i) {
                    const int e = g0 + i;
                    const Pt p = point<COL>(e, vb);
                    const auto& o = e < PF ? ah.op[e < PF ? e : 0] : rest[e >= PF ? e - PF : 0];
                    if constexpr (MAP) v[e] = f.apply(p, LOADG ? make_float2(0.f, 0.f) : v[e], o);
                    else f.apply(p, v[e], o);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            return;
        }
#pragma unroll
        for (int g0 = 0; g0 < E; g0 += G) {
            decltype(f.load(Pt{})) op[G];
            if (g0 >= PF) {
#pragma unroll
                for (int i = 0; i < G; ++i) op[i] = f.load(point<COL>(g0 + i, vb));
                __builtin_amdgcn_sched_barrier(0);
            } else {
#pragma unroll
                for (int i = 0; i < G; ++i) op[i] = ah.op[g0 + i];
            }
#pragma unroll
            for (int i = 0; i < G; ++i) {
                const Pt p = point<COL>(g0 + i, vb);
                if constexpr (MAP) v[g0 + i] = f.apply(p, LOADG ? make_float2(0.f, 0.f) : v[g0 + i], op[i]);
                else f.apply(p, v[g0 + i], op[i]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // the post functor of a transform: its first PF operands are loaded before the last pass
    template <bool COL, class F, class PASS> __device__ __forceinline__ void last_pass_and_post(const F& f, PASS pass) {
        if constexpr (IsStaged<F>::value && PrefetchOf<F>::value > 0 && AHEAD) {
            Ahead<F> ah;
            load_ahead<COL>(f, ah);
            pass();
            run_staged<COL, true, false, PrefetchOf<F>::value>(f, ah);
        } else {
            pass();
            run_elems<COL, true, false>(f);
        }
    }

    template <bool LOADG = false, class Pre = NoOp, class Post = NoOp>
    __device__ __forceinline__ void fwd(const Pre& pre = Pre(), const Post& post = Post()) {
        opaque();
        run_elems<true, true, LOADG>(pre);
        pass_dit<false>();
        transpose_c_to_r();
        last_pass_and_post<false>(post, [&] { pass_dit<false>(); });
    }

    template <bool LOADG = false, class Pre = NoOp, class Post = NoOp>
    __device__ __forceinline__ void inv(const Pre& pre = Pre(), const Post& post = Post()) {
        opaque();
        run_elems<false, true, LOADG>(pre);
        pass_dif<true>();
        transpose_r_to_c();
        last_pass_and_post<true>(post, [&] { pass_dif<true>(); });
    }

    template <class F> __device__ __forceinline__ void each_R(const F& f) {
        opaque();
        run_elems<false, false, false>(f);
    }
    template <class F> __device__ __forceinline__ void each_C(const F& f) {
        opaque();
        run_elems<true, false, false>(f);
    }
};

}  // namespace mset
