// fft_reg.hpp -- register-resident 2D FFT of a 128 x 128 complex plane for CDNA4 (gfx950).
//
// The workgroup has 512 threads.  Four consecutive lanes (a quad) own one line of the plane (a
// column in the "C" orientation, a row in the "R" orientation), 32 elements each, in registers.  A
// 1D transform of a 128-point line is a 32-point DFT in registers, a twiddle multiply, and a 4-point
// DFT across the quad done with DPP lane swaps (quad_perm xor 1 / xor 2, no LDS, no barrier).  The
// twiddles W_128^(b(h) k) come from the LDS table (the four lanes of a quad read four entries, all
// quads the same four: a broadcast):
//
//   decimation in time (forward):  lane h holds x[b(h) + 4m]   ->  lane h holds X[32 h + k]
//   decimation in frequency (inv): lane h holds x[32 h + m]    ->  lane h holds X[b(h) + 4 k]
//
// with b the 2-bit bit reversal.  The two layouts chain: a forward transform ends with blocked lines
// and the inverse starts from them, and the inverse ends with decimated lines, which is what the next
// forward transform consumes, so consecutive transforms never reshuffle registers.  The only LDS
// traffic of a 2D transform is one transpose between its column and row passes (C <-> R), through a
// 135 KB buffer whose swizzle A(y, x) = 132 y + x + 4 (y >> 5) keeps every 16-lane access on 16
// distinct 8-byte bank groups for all four access patterns.
//
// Interface (shared with LdsEngine in fft_lds.hpp): fwd(pre, post) goes C -> R and inv(pre, post)
// goes R -> C; pre / post see (y, x, value) for every element the thread owns; each_C / each_R visit
// the owned elements after an inverse / forward transform.  In C, consecutive lanes hold consecutive
// x (coalesced for row-major [y][x] arrays); in R, consecutive lanes hold consecutive y (coalesced for
// arrays stored [x][y]).
#pragma once
#include <hip/hip_runtime.h>
#include "fft_lds.hpp"

namespace mset {

template <int CTRL> __device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
template <int CTRL> __device__ __forceinline__ float2 dpp_c(float2 v) {
    return make_float2(dpp_f<CTRL>(v.x), dpp_f<CTRL>(v.y));
}
constexpr int DPP_XOR1 = 0xB1;   // quad_perm [1, 0, 3, 2]
constexpr int DPP_XOR2 = 0x4E;   // quad_perm [2, 3, 0, 1]

// lane-sign combine: lanes with the bit clear keep own + partner, the others partner - own
__device__ __forceinline__ float2 bfly(float2 own, float2 partner, float sgn) {
    return make_float2(fmaf(sgn, own.x, partner.x), fmaf(sgn, own.y, partner.y));
}

template <int N> struct RegEngine;

template <> struct RegEngine<128> {
    static constexpr int N = 128, E = 32, NT = 512, SP = 132;
    static constexpr int LDS_ELEMS = 127 * SP + 127 + 4 * 3 + 1;
    float2* s;
    const float2* tw;  // W_128^k table in LDS; lanes of one b(h) read the same entry (broadcast)
    float2 v[E];
    int line, h, b;
    float s1, s2;      // lane signs for the xor-1 and xor-2 butterflies
    bool lane3;

    __device__ __forceinline__ void init(float2* lds, const float2* lds_tw) {
        s = lds;
        tw = lds_tw;
        const int t = threadIdx.x;
        line = t >> 2;
        h = t & 3;
        b = ((h & 1) << 1) | (h >> 1);
        s1 = (h & 1) ? -1.f : 1.f;
        s2 = (h & 2) ? -1.f : 1.f;
        lane3 = h == 3;
    }

    static __device__ __forceinline__ int A(int y, int x) { return y * SP + x + 4 * (y >> 5); }

    // The element functors load global operands; unrolled over 32 elements the scheduler would hoist
    // every load ahead of its use and run out of registers.  A scheduling fence after every G
    // elements bounds the work in flight to G elements' worth.
    template <int G = 4> static __device__ __forceinline__ void group_fence(int i) {
        if ((i & (G - 1)) == G - 1) __builtin_amdgcn_sched_barrier(0);
    }

    // Per-element addresses are cheap affine functions of (line, h, b).  Left visible, the compiler
    // hoists all of them out of the plane and position loops and keeps hundreds of registers live;
    // routing the coordinates through an empty asm at every transform makes them opaque, so the
    // addresses are recomputed where they are used.
    __device__ __forceinline__ void opaque() {
        asm volatile("" : "+v"(line), "+v"(h), "+v"(b));
    }

    // lanes hold z_{b(h)}; afterwards lane h holds sum_j z_j W4^(j h)
    template <bool INV> __device__ __forceinline__ float2 xdit(float2 z) const {
        z = bfly(z, dpp_c<DPP_XOR1>(z), s1);
        z = lane3 ? mul_mi<INV>(z) : z;
        return bfly(z, dpp_c<DPP_XOR2>(z), s2);
    }
    // lanes hold z_h; afterwards lane h holds sum_j z_j W4^(j b(h))
    template <bool INV> __device__ __forceinline__ float2 xdif(float2 z) const {
        z = bfly(z, dpp_c<DPP_XOR2>(z), s2);
        z = lane3 ? mul_mi<INV>(z) : z;
        return bfly(z, dpp_c<DPP_XOR1>(z), s1);
    }

    template <bool INV> __device__ __forceinline__ void pass_dit() {
        DFT<E, INV>::run(v);
#pragma unroll
        for (int k = 1; k < E; ++k) {
            const float2 w = tw[b * k];
            v[k] = INV ? cmulc(v[k], w) : cmul(v[k], w);
            group_fence<8>(k);
        }
#pragma unroll
        for (int k = 0; k < E; ++k) {
            v[k] = xdit<INV>(v[k]);
            group_fence<8>(k);
        }
    }
    template <bool INV> __device__ __forceinline__ void pass_dif() {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            v[m] = xdif<INV>(v[m]);
            group_fence<8>(m);
        }
#pragma unroll
        for (int m = 1; m < E; ++m) {
            const float2 w = tw[b * m];
            v[m] = INV ? cmulc(v[m], w) : cmul(v[m], w);
            group_fence<8>(m);
        }
        DFT<E, INV>::run(v);
    }

    // column-blocked (y = 32 h + k of column `line`) -> row-decimated (x = b + 4 m of row `line`)
    __device__ __forceinline__ void transpose_c_to_r() {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < E; ++k) s[A(32 * h + k, line)] = v[k];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = s[A(line, b + 4 * m)];
    }
    // row-decimated (x = b + 4 k of row `line`) -> column-blocked (y = 32 h + m of column `line`)
    __device__ __forceinline__ void transpose_r_to_c() {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < E; ++k) s[A(line, b + 4 * k)] = v[k];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = s[A(32 * h + m, line)];
    }

    // Visit the 32 owned elements.  Column-decimated coordinates (COL): element m is (b + 4 m, line);
    // row-blocked ones: element k is (line, 32 h + k).  MAP: the functor returns the new value (LOADG:
    // it is handed zero instead of the current one); otherwise it sees the element by reference.
    template <bool COL, bool MAP, bool LOADG, class F> __device__ __forceinline__ void run_elems(const F& f) {
        auto yx = [&](int i, int& y, int& x) {
            if (COL) { y = b + 4 * i; x = line; }
            else { y = line; x = 32 * h + i; }
        };
        if constexpr (IsStaged<F>::value) {
            constexpr int G = F::G;
            static_assert(E % G == 0, "group size must divide 32");
#pragma unroll
            for (int g0 = 0; g0 < E; g0 += G) {
                decltype(f.load(0, 0)) op[G];
#pragma unroll
                for (int i = 0; i < G; ++i) {
                    int y, x;
                    yx(g0 + i, y, x);
                    op[i] = f.load(y, x);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int i = 0; i < G; ++i) {
                    int y, x;
                    yx(g0 + i, y, x);
                    if constexpr (MAP) v[g0 + i] = f.apply(y, x, LOADG ? make_float2(0.f, 0.f) : v[g0 + i], op[i]);
                    else f.apply(y, x, v[g0 + i], op[i]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
#pragma unroll
            for (int i = 0; i < E; ++i) {
                int y, x;
                yx(i, y, x);
                if constexpr (MAP) v[i] = f(y, x, LOADG ? make_float2(0.f, 0.f) : v[i]);
                else f(y, x, v[i]);
                group_fence(i);
            }
        }
    }

    template <bool LOADG = false, class Pre = NoOp, class Post = NoOp>
    __device__ __forceinline__ void fwd(const Pre& pre = Pre(), const Post& post = Post()) {
        opaque();
        run_elems<true, true, LOADG>(pre);
        pass_dit<false>();
        transpose_c_to_r();
        pass_dit<false>();
        run_elems<false, true, false>(post);
    }

    template <bool LOADG = false, class Pre = NoOp, class Post = NoOp>
    __device__ __forceinline__ void inv(const Pre& pre = Pre(), const Post& post = Post()) {
        opaque();
        run_elems<false, true, LOADG>(pre);
        pass_dif<true>();
        transpose_r_to_c();
        pass_dif<true>();
        run_elems<true, true, false>(post);
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
