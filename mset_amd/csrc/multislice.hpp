// multislice.hpp -- fused multislice forward / adjoint kernels (CDNA4, gfx950): the position routines
// and the kernel templates.  The launchers instantiate them in three translation units (multislice.hip:
// sequential, forward-only and plane-FFT kernels; batch_store.hip: the batched gradient with stored waves;
// batch_variants.hip: with recomputed or checkpointed waves), so that the build compiles them in parallel.
//
// One workgroup owns one probe position at a time and keeps the N x N wave in LDS for the whole
// forward and adjoint sweep.  Everything the reference does per position with separate launches
// (src/CUDA/mset.cu:107-333: probe shift, per-plane x t / FFT / x p / IFFT, exit FFT, residual,
// per-plane adjoint and gradient, probe and position terms) is one pass through this routine; the
// element-wise steps ride on the FFT stages' loads and stores (fft_lds.hpp).
//
// Three kinds of launch share it:
//   K_PLAIN  one position per launch, sequential semantics of src/Run_MSET_recon.m:134-158 and
//            src/CUDA/mset.cu:107-333: the transmission planes are recomputed from V at the start, and
//            V, the probe and the position are updated in place at the end, before the next launch.
//   K_BATCH  many positions per launch at a fixed (V, probe): a persistent grid of workgroups, each
//            accumulating its positions' object gradient into a private slab (L planes) and its
//            probe-gradient term into a private Fourier-domain plane; slabs are summed afterwards in
//            a fixed order, so results are bitwise repeatable and no atomics are used.
//   K_FWD    forward + loss only (run_error_mset, src/CUDA/mset.cu:557-696), many positions.
//
// Conventions: Fourier data is unshifted (DC at [0][0]); the propagator and the measured patterns are
// ifftshifted once at upload (src/CUDA/mex_main.cu:80, mset.cu:208).  Real-space data keeps the
// MATLAB layout; the multislice operator commutes with circular shifts, so this equals MATLAB's
// fftshift(fftn(ifftshift(.))) chain (STEP02_FORWARD_4D.m:32-36).  The sub-pixel shift follows
// Func_generate_probe_wave_scan_sub_pixel.m:4-27 (grid -n/2..n/2-1, so the Nyquist bin uses -n/2).
//
// Global layouts (chosen so that every fused access is coalesced, see fft_lds.hpp): row-major [y][x]
// for V, transmission planes, saved waves, gradient slabs and the probe; transposed [x][y] for the
// arrays read in the row pass that ends a forward transform -- the propagator, sqrt(M), fftn(probe)
// and the Fourier-domain incident wave.
#pragma once
#include <hip/hip_runtime.h>
#include "fft_lds.hpp"
#include "fft_reg.hpp"
#include "fft_tiled.hpp"
#include "kernels.hpp"

namespace mset {

// plane engine and workgroup size per plane size: register-resident at 128, LDS stages elsewhere
// Engine and workgroup shape per plane size.  NT threads; MINB: waves per SIMD the register budget must allow
// (HIP launch bounds).  SH = 0: the shape every kernel uses; SH = 1: the alternative shape of the batched
// kernel, a candidate for the autotuner (MSET_ENGINE_ALT_SHAPE): twice the threads per position of the LDS
// engine up to N = 64, 1024 threads at 96, at 128 the register engine with a half-size transpose buffer and a
// 128-register budget (two workgroups per CU instead of one; it spills about 400 registers), and tiles of 16
// lines with 256 threads (three workgroups per CU instead of two) at 256.
template <int N, int SH = 0> struct Cfg;
template <> struct Cfg<16> { static constexpr int NT = 64, MINB = 1;  using Eng = LdsEngine<16, 64>; };
template <> struct Cfg<24> { static constexpr int NT = 64, MINB = 1;  using Eng = LdsEngine<24, 64>; };
template <> struct Cfg<32> { static constexpr int NT = 64, MINB = 1;  using Eng = LdsEngine<32, 64>; };
template <> struct Cfg<48> { static constexpr int NT = 192, MINB = 1; using Eng = LdsEngine<48, 192>; };
template <> struct Cfg<64> { static constexpr int NT = 256, MINB = 3; using Eng = LdsEngine<64, 256>; };
template <> struct Cfg<96> { static constexpr int NT = 576, MINB = 1; using Eng = LdsEngine<96, 576>; };
template <> struct Cfg<128> { static constexpr int NT = 512, MINB = 1; using Eng = RegEngine<128>; };
template <> struct Cfg<256> { static constexpr int NT = 512, MINB = 4; using Eng = TiledEngine<256, 512>; };
template <> struct Cfg<16, 1> { static constexpr int NT = 128, MINB = 1; using Eng = LdsEngine<16, 128>; };
template <> struct Cfg<24, 1> { static constexpr int NT = 128, MINB = 1; using Eng = LdsEngine<24, 128>; };
template <> struct Cfg<32, 1> { static constexpr int NT = 128, MINB = 1; using Eng = LdsEngine<32, 128>; };
template <> struct Cfg<48, 1> { static constexpr int NT = 384, MINB = 1; using Eng = LdsEngine<48, 384>; };
template <> struct Cfg<64, 1> { static constexpr int NT = 512, MINB = 1; using Eng = LdsEngine<64, 512>; };
template <> struct Cfg<96, 1> { static constexpr int NT = 1024, MINB = 1; using Eng = LdsEngine<96, 1024>; };
template <> struct Cfg<128, 1> { static constexpr int NT = 512, MINB = 4; using Eng = RegEngine<128, true>; };
template <> struct Cfg<256, 1> { static constexpr int NT = 256, MINB = 3; using Eng = TiledEngine<256, 256>; };
// launches of at most one position per CU (a tilt's last mini-batch): 64-line tiles, one workgroup of 16 waves
// per CU, twice the default workgroup's loads in flight
template <> struct Cfg<256, 2> { static constexpr int NT = 1024, MINB = 4; using Eng = TiledEngine<256, 1024>; };


enum { K_PLAIN = 0, K_BATCH = 1, K_FWD = 2 };

// Sum of one double per thread over the workgroup; every thread gets the total.
template <int NT> __device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < (NT + 63) / 64; ++i) s += red[i];
    return s;
}

// grid value of the sub-pixel shift ramp at unshifted frequency index j (Nyquist -> -N/2)
template <int N> __device__ __forceinline__ double shift_grid(int j) {
    return (double)((j + N / 2) % N) - (double)(N / 2);
}

// global operands of one element of the adjoint step: transmission, saved wave, gradient slab
struct AdjOp {
    float2 t, w;
    float g;
};
// the same with the transmission as its phase (separable routine)
struct AdjPh {
    float ph;
    float2 w;
    float g;
};


// slot: the workgroup's scratch (waves, plane buffers); slab: the gradient slab the position adds to;
// first: the slab's first position of the launch, which writes the slab instead of adding to it (so the
// slabs need no clearing launch).  PFAC: the propagator is applied as the product of its 1D factors
// pxs[kx] pys[ky] (in LDS) instead of being read from the 2D array (separable propagator, below 256).
template <int N, int NT, int KIND, int WV = WAVES_STORE, bool PFAC = false, class Eng>
__device__ __forceinline__ void process_position(const PosArgs& A, int k, int slot, int slab, bool first, Eng& E,
                                                 float2* phx, float2* phy, double* red,
                                                 const float2* pxs = nullptr, const float2* pys = nullptr) {
    constexpr int N2 = N * N;
    const int t = threadIdx.x;
    const int L = A.L;
    const float inv_n2 = 1.0f / (float)N2;
    // Per-slot scratch.  Store: planes 0..L hold the waves psi_a [y][x], plane L+1 the Fourier psi0 [x][y].
    // Recompute (batched kernel only): plane 0 psi0, plane 1 Fourier psi0, planes 2 and 3 the wave on
    // either side of the current adjoint plane, plane 4 the parked adjoint wave.
    // Checkpoint (batched kernel only): plane 0 psi0, plane 1 Fourier psi0, planes 1 + s the checkpoints
    // psi_(s K), s = 1 .. L / K, planes 1 + L / K + m the waves psi_(s K + m), m = 1 .. K - 1, of the segment
    // s the adjoint is in, and the last plane the parked adjoint wave.
    constexpr bool recompute = WV == WAVES_RECOMPUTE && KIND == K_BATCH;
    constexpr bool ckpt = WV == WAVES_CHECKPOINT && KIND == K_BATCH;
    const int K = ckpt ? A.ckseg : 1, SL = L / K;
    float2* save = A.save + (size_t)slot * (recompute ? RECOMPUTE_PLANES : ckpt ? checkpoint_planes(L, K) : L + 2) * N2;
    float2* F0 = save + (size_t)(recompute || ckpt ? 1 : L + 1) * N2;
    // checkpoint variant: where wave psi_i lives (a checkpoint, psi0, or the segment buffer)
    auto ck_wave = [&](int i) -> float2* {
        const int m = i % K;
        return save + (size_t)(m != 0 ? 1 + SL + m : i == 0 ? 0 : 1 + i / K) * N2;
    };
    E.bind(A.planes ? A.planes + (size_t)slot * 2 * N2 : nullptr);
    const bool want_pos = (A.want & WANT_POS) != 0;
    const bool want_probe = (A.want & WANT_PROBE) != 0;
    const bool exact = A.grad_mode == GRAD_EXACT;
    const float2* pT = A.prop;                              // propagator, [x][y]

    // ---- sub-pixel shift factors, separable: Phi(y, x) = phy[y] * phx[x]  (double-precision phase)
    const double row = (double)A.xlist[k], col = (double)A.ylist[k];
    const double dxs = row - (double)A.cent, dys = col - (double)A.cent;
    for (int j = t; j < 2 * N; j += NT) {
        const int jj = j < N ? j : j - N;
        const double d = j < N ? dxs : dys;
        const double a = -2.0 * M_PI * d * shift_grid<N>(jj) / (double)N;
        double sn, cs;
        sincos(a, &sn, &cs);
        (j < N ? phx : phy)[jj] = make_float2((float)cs, (float)sn);
    }
    __syncthreads();
    auto phase = [&](int y, int x) { return cmul(phy[y], phx[x]); };

    // The sequential kernel reads its transmission planes from A.trans like the batched ones: the host
    // recomputes them from the current volume with a full-device launch before every position, and applies
    // the object step with another after it (one workgroup streaming the whole volume twice per position
    // took 28 % of the position's time).

    // buffer views (hardware bounds-checked); p.e indexes the layout the engine's orientation streams
    const BufC pTb(pT, N2);                                 // propagator, [x][y]
    const BufC sv0(save, 2 * N2);                           // psi0 and (single slice) the exit wave, [y][x]
    const BufC F0b(F0, N2);                                 // Fourier psi0, [x][y]
    const BufC tr0(A.trans, N2);                            // transmission of plane 0, [y][x]

    // ---- incident wave psi0 = ifftn(fftn(probe) .* Phi), stored, times t_0.  Each inverse transform of
    // the forward sweep ends with the next plane's transmission, whose operands are loaded ahead of the
    // last pass: the loads are then issued before the wave's stores and never wait behind them.
    const bool multi = !A.sset && L > 0;
    // An engine whose plane lives in global memory fuses the passes that meet between transforms
    // (row_step, col_step): from here to the end of the adjoint the plane alternates between the
    // half-forward and half-inverse states.
    constexpr bool FUSE = Eng::FUSED;
    auto psi0_post = staged<16, 32>([&](const Pt& p) { return tr0[p.e]; }, [&](const Pt& p, float2 v, float2 o) {
        float2 w = cscale(v, inv_n2);
        if (KIND != K_FWD) sv0.put(p.e, w);
        w = cmul(w, o);
        // single slice: the exit wave is psi0 .* t, no propagation
        if (A.sset && KIND != K_FWD) sv0.put(Idx{p.e.v, p.e.s + (unsigned)N2}, w);
        return w;
    });
    if constexpr (KIND == K_PLAIN) {
        const BufC pr(A.probe, N2);
        E.template fwd<true>(staged<16>([&](const Pt& p) { return pr[p.e]; },
                                        [&](const Pt&, float2, float2 o) { return o; }),
                             [&](const Pt& p, float2 v) {
                                 const float2 f = cmul(v, phase(p.y, p.x));
                                 F0b.put(p.e, f);
                                 return f;
                             });
        if constexpr (FUSE) {
            E.row_first(NoOp());
            E.col_step(psi0_post);
        } else {
            E.inv(NoOp(), psi0_post);
        }
    } else {
        const BufC pf(A.probeF, N2);                        // fftn(probe), [x][y]
        auto spectrum = staged<16>([&](const Pt& p) { return pf[p.e]; }, [&](const Pt& p, float2, float2 o) {
            const float2 f = cmul(o, phase(p.y, p.x));
            if (KIND == K_BATCH) F0b.put(p.e, f);
            return f;
        });
        if constexpr (FUSE) {
            E.template row_first<true>(spectrum);
            E.col_step(psi0_post);
        } else {
            E.template inv<true>(spectrum, psi0_post);
        }
    }

    // ---- forward sweep: psi_a = ifftn(fftn(psi_{a-1} .* t_a) .* p); the engine enters plane a holding
    // psi_{a-1} .* t_a
    auto pfac = [&](const Pt& p) { return cmul(pxs[p.x], pys[p.y]); };
    for (int a = 0; a < (multi ? L : 0); ++a) {
        if constexpr (PFAC) {
            auto byp = [&](const Pt& p, float2 v) { return cmul(v, pfac(p)); };
            if constexpr (FUSE) E.row_step(byp);
            else E.fwd(NoOp(), byp);
        } else {
            auto byp = staged<16, 32>([&](const Pt& p) { return pTb[p.e]; }, [&](const Pt&, float2 v, float2 o) { return cmul(v, o); });
            if constexpr (FUSE) E.row_step(byp);
            else E.fwd(NoOp(), byp);
        }
        const bool next = a + 1 < L;
        // recompute keeps only the exit wave psi_L (plane 2), the start of the adjoint's rebuild; checkpoint
        // keeps psi_(a+1) if it is a checkpoint or in the last segment, which the adjoint meets first
        const bool keep = KIND != K_FWD && (recompute ? !next : ckpt ? ((a + 1) % K == 0 || (a + 1) / K == SL) : true);
        const BufC sv(recompute ? save + 2 * (size_t)N2 : ckpt ? ck_wave(a + 1) : save + (size_t)(a + 1) * N2, N2);
        // the last plane has no successor: an empty view reads zeros and the product is skipped
        const BufC tn(A.trans + (size_t)(next ? a + 1 : 0) * N2, next ? N2 : 0u);
        auto post = staged<16, 32>([&](const Pt& p) { return tn[p.e]; }, [&](const Pt& p, float2 v, float2 o) {
            const float2 w = cscale(v, inv_n2);
            if (keep) sv.put(p.e, w);
            return next ? cmul(w, o) : w;
        });
        if constexpr (FUSE) E.col_step(post);
        else E.inv(NoOp(), post);
    }

    // ---- exit wave, intensity, amplitude residual (STEP02_FORWARD_4D.m:36-49), r .* conj(p) fused
    double err = 0.0;
    {
        const BufF MT(A.sqrtM + (size_t)k * N2, N2);        // sqrt(M), [x][y]
        float* Iout = A.inten ? A.inten + (size_t)k * N2 : nullptr;
        auto exitf = staged<16, 16>(
                          [&](const Pt& p) {
                              const float m = MT[p.e];
                              if (KIND == K_FWD) return make_float3(m, 0.f, 0.f);
                              if (A.sset) return make_float3(m, 1.f, 0.f);     // no back-propagator
                              if constexpr (PFAC) return make_float3(m, 0.f, 0.f);
                              const float2 q = pTb[p.e];
                              return make_float3(m, q.x, q.y);
                          },
                          [&](const Pt& p, float2 v, float3 o) {
                              const float I = v.x * v.x + v.y * v.y;
                              if (KIND == K_FWD && Iout) Iout[((p.y + N / 2) % N) * N + (p.x + N / 2) % N] = I;
                              const float m = o.x;
                              // hardware square root and reciprocal square root (1 ulp)
                              const float d = m - __builtin_amdgcn_sqrtf(I);
                              err += (double)(d * d);
                              if (KIND == K_FWD) return v;
                              const float f = 1.0f - m * __builtin_amdgcn_rsqf(I + 1e-30f);
                              return cmulc(make_float2(v.x * f, v.y * f), PFAC && !A.sset ? pfac(p) : make_float2(o.y, o.z));
                          });
        if constexpr (FUSE) {
            // the adjoint's first row pass rides along (none for the error pass)
            if constexpr (KIND == K_FWD) E.row_last(exitf);
            else E.row_step(exitf);
        } else {
            E.fwd(NoOp(), exitf);
        }
    }
    err = block_sum<NT>(err, red);
    if (t == 0) A.err_pos[k - A.obase] = err;
    if constexpr (KIND == K_FWD) return;

    // ---- adjoint sweep (STEP03_BACK.m:12-24, mset.cu:232-264)
    // Recompute: the forward sweep is inverted plane by plane, psi_a = conj(t_a) .* ifftn(fftn(psi_{a+1})
    // .* conj(p)), exact because the propagator and the transmission have unit modulus (the host only
    // selects this variant for a unit-modulus propagator).  cur holds psi_{a+1} at the top of plane a.
    // Checkpoint: the adjoint of plane a pairs psi_i (i = a + 1, or a in the exact mode), i descending; on
    // entering segment s = i / K below the one held, the segment's waves psi_(s K + 1) .. psi_(s K + K - 1)
    // are rebuilt from its checkpoint by the forward steps of planes s K .. s K + K - 2, the same arithmetic
    // as the forward sweep's.  cseg is the segment held (the last one, stored by the forward sweep).
    float2* cur = save + 2 * (size_t)N2;
    float2* prev = save + 3 * (size_t)N2;
    const BufC park(save + (size_t)(ckpt ? checkpoint_planes(L, K) - 1 : 4) * N2, N2);
    int cseg = SL;
    // checkpoint: the forward steps of segment cseg from its first wave, each wave stored
    auto rebuild_segment = [&]() {
        const int i0 = cseg * K;
        const BufC cb(ck_wave(i0), N2), tb0(A.trans + (size_t)i0 * N2, N2);
        auto enter = staged<16>([&](const Pt& p) { const float2 w = cb[p.e], tt = tb0[p.e]; return make_float4(w.x, w.y, tt.x, tt.y); },
                                [&](const Pt&, float2, float4 o) { return cmul(make_float2(o.x, o.y), make_float2(o.z, o.w)); });
        auto byp = staged<16, 32>([&](const Pt& p) { return pTb[p.e]; }, [&](const Pt&, float2 v, float2 o) { return cmul(v, o); });
        if constexpr (FUSE) E.template col_first<true>(enter);
        else E.template fwd<true>(enter, byp);
        for (int a2 = i0; a2 < i0 + K - 1; ++a2) {
            const bool more = a2 + 2 < i0 + K;                // another plane of the segment follows
            const BufC sb(ck_wave(a2 + 1), N2);
            const BufC tn(A.trans + (size_t)(more ? a2 + 1 : 0) * N2, more ? N2 : 0u);
            auto post = staged<16, 32>([&](const Pt& p) { return tn[p.e]; }, [&](const Pt& p, float2 v, float2 o) {
                const float2 w = cscale(v, inv_n2);
                sb.put(p.e, w);
                return more ? cmul(w, o) : w;
            });
            if constexpr (FUSE) {
                E.row_step(byp);
                if (more) E.col_step(post);
                else E.col_last(post);
            } else {
                E.inv(NoOp(), post);
                if (more) E.fwd(NoOp(), byp);
            }
        }
    };
    for (int a = L - 1; a >= 0; --a) {
        const BufC ta(A.trans + (size_t)a * N2, N2);
        // the single-slice operator pairs the incident wave (STEP03_BACK_SSET.m:13-14)
        const float2* wsp = save + (size_t)((exact || A.sset) ? a : a + 1) * N2;
        bool rebuild = recompute && a > 0;
        if (recompute) wsp = exact ? (a > 0 ? prev : save) : cur;
        if constexpr (ckpt) {
            const int i = exact ? a : a + 1;
            rebuild = i / K != cseg;
            if (rebuild) cseg = i / K;
            wsp = ck_wave(i);
        }
        if constexpr (ckpt) {
            if (rebuild) {
                if constexpr (FUSE) {
                    // the adjoint wave (half-inverse state) stays in its buffer; the rebuild runs on the park plane
                    float2* held = E.swap_rs(save + (size_t)(checkpoint_planes(L, K) - 1) * N2);
                    rebuild_segment();
                    E.swap_rs(held);
                } else {
                    E.each_R([&](const Pt& p, float2& v) { park.put(p.e, v); });
                    rebuild_segment();
                }
            }
        } else if constexpr (KIND == K_BATCH && FUSE) {
            if (rebuild) {
                // the adjoint wave (half-inverse state) stays in its buffer while the engine rebuilds psi_a
                // into prev on the park plane: three fused passes, no copies
                float2* held = E.swap_rs(save + 4 * (size_t)N2);
                const BufC cb(cur, N2), pb(prev, N2);
                E.template col_first<true>(staged<16>([&](const Pt& p) { return cb[p.e]; }, [&](const Pt&, float2, float2 o) { return o; }));
                E.row_step(staged<16, 32>([&](const Pt& p) { return pTb[p.e]; },
                                          [&](const Pt&, float2 v, float2 o) { return cmulc(v, o); }));
                E.col_last(staged<16, 32>([&](const Pt& p) { return ta[p.e]; }, [&](const Pt& p, float2 v, float2 o) {
                    const float2 w = cmulc(cscale(v, inv_n2), o);
                    pb.put(p.e, w);
                    return w;
                }));
                E.swap_rs(held);
            }
        } else if constexpr (KIND == K_BATCH) {
            if (rebuild) {
                // park the adjoint wave (Fourier domain), rebuild psi_a into prev
                E.each_R([&](const Pt& p, float2& v) { park.put(p.e, v); });
                const BufC cb(cur, N2), pb(prev, N2);
                E.template fwd<true>(staged<16>([&](const Pt& p) { return cb[p.e]; }, [&](const Pt&, float2, float2 o) { return o; }),
                                     staged<16, 32>([&](const Pt& p) { return pTb[p.e]; },
                                                    [&](const Pt&, float2 v, float2 o) { return cmulc(v, o); }));
                E.inv(NoOp(), staged<16, 32>([&](const Pt& p) { return ta[p.e]; }, [&](const Pt& p, float2 v, float2 o) {
                    const float2 w = cmulc(cscale(v, inv_n2), o);
                    pb.put(p.e, w);
                    return w;
                }));
            }
        }
        const BufC ws(wsp, N2);
        const BufF g(A.gslab + ((size_t)slab * L + a) * N2, N2);
        const BufF gin(A.gslab + ((size_t)slab * L + a) * N2, first ? 0u : N2);   // empty view: reads 0
        auto adjoint = [&](float2 v, float2 t, float2 w, float& gr) {
            const float2 b = cscale(v, inv_n2);
            const float2 tc = cconj(t);
            const float2 q = cmul(tc, cmulc(b, w));
            gr = -q.y;                                        // Re(1i * conj(t) * conj(psi) * b)
            return cmul(b, tc);
        };
        if constexpr (KIND == K_PLAIN) {
            // the gradient plane of this position (applied to the bin after the sweep)
            auto post = staged<16, 16>(
                [&](const Pt& p) {
                    const float2 tt = ta[p.e], ww = ws[p.e];
                    return make_float4(tt.x, tt.y, ww.x, ww.y);
                },
                [&](const Pt& p, float2 v, float4 o) {
                    float gr;
                    const float2 r = adjoint(v, make_float2(o.x, o.y), make_float2(o.z, o.w), gr);
                    g.put(p.e, gr);
                    return r;
                });
            if constexpr (FUSE) E.col_step(post);
            else E.inv(NoOp(), post);
        } else {
            // accumulated into this workgroup's slab
            auto post = staged<8, 16, true>([&](const Pt& p) { return AdjOp{ta[p.e], ws[p.e], gin[p.e]}; },
                                            [&](const Pt& p, float2 v, const AdjOp& o) {
                                                float gr;
                                                const float2 r = adjoint(v, o.t, o.w, gr);
                                                g.put(p.e, o.g + gr);
                                                return r;
                                            });
            if constexpr (FUSE) E.col_step(post);
            else if (rebuild)   // the parked adjoint wave comes back as the inverse transform's input
                E.template inv<true>(staged<16>([&](const Pt& p) { return park[p.e]; }, [&](const Pt&, float2, float2 o) { return o; }),
                                     post);
            else E.inv(NoOp(), post);
        }
        if (a > 0) {
            if constexpr (PFAC) {
                auto bypc = [&](const Pt& p, float2 v) { return cmulc(v, pfac(p)); };
                if constexpr (FUSE) E.row_step(bypc);
                else E.fwd(NoOp(), bypc);
            } else {
                auto bypc = staged<16, 32>([&](const Pt& p) { return pTb[p.e]; },
                                           [&](const Pt&, float2 v, float2 o) { return cmulc(v, o); });
                if constexpr (FUSE) E.row_step(bypc);
                else E.fwd(NoOp(), bypc);
            }
        } else {
            if constexpr (FUSE) E.row_last(NoOp());
            else E.fwd();
        }
        if (recompute && rebuild) {
            float2* t2 = cur;
            cur = prev;
            prev = t2;
        }
    }
    // the engine now holds fftn(back_wave_f) (Fourier domain, R state)

    // ---- probe term in the Fourier domain: shift(bw, -d) = ifftn(fftn(bw) .* conj(Phi)); slab [x][y]
    if constexpr (KIND == K_BATCH) {
        if (want_probe) {
            const BufC ps(A.pslab + (size_t)slab * N2, N2);
            const BufC psin(A.pslab + (size_t)slab * N2, first ? 0u : N2);   // empty view: reads 0
            E.each_R(staged<16>([&](const Pt& p) { return psin[p.e]; },
                                [&](const Pt& p, float2& v, float2 o) { ps.put(p.e, cadd(o, cmulc(v, phase(p.y, p.x)))); }));
        }
    }

    // ---- position terms (STEP03_BACK.m:89-92)
    double dr = 0.0, dc = 0.0;
    bool have_bw = false;
    if (want_pos) {
        if (exact) {
            // Parseval: sum_x bw conj(-dpsi0/drow) = (1/N^2) sum_q B(q) conj(Psi0(q)) (-2 pi i g/N)
            E.each_R(staged<16>([&](const Pt& p) { return F0b[p.e]; }, [&](const Pt& p, float2& v, float2 o) {
                const float2 c = cmulc(v, o);                  // B conj(Psi0)
                // Re(c * (-i) * 2 pi g / N) = Im(c) * 2 pi g / N
                dr += (double)c.y * shift_grid<N>(p.x);
                dc += (double)c.y * shift_grid<N>(p.y);
            }));
            const double f = 2.0 * M_PI / (double)N / (double)N2;
            dr = block_sum<NT>(dr, red) * f;
            dc = block_sum<NT>(dc, red) * f;
        } else {
            // back_wave_f = ifftn(B); MATLAB gradient() of conj(psi0), one-sided at the edges
            E.inv(NoOp(), [&](const Pt&, float2 v) { return cscale(v, inv_n2); });
            have_bw = true;
            const float2* p0 = save;
            E.each_C([&](const Pt& p, float2& bw) {
                const int y = p.y, x = p.x;
                // clamped neighbours: the edge rows / columns become the one-sided differences
                const int xp = min(x + 1, N - 1), xm = max(x - 1, 0);
                const int yp = min(y + 1, N - 1), ym = max(y - 1, 0);
                const float2 gx = cscale(csub(p0[y * N + xp], p0[y * N + xm]), xp - xm == 2 ? 0.5f : 1.0f);
                const float2 gy = cscale(csub(p0[yp * N + x], p0[ym * N + x]), yp - ym == 2 ? 0.5f : 1.0f);
                // Re(conj(g) * bw)
                dr += (double)(gx.x * bw.x + gx.y * bw.y);
                dc += (double)(gy.x * bw.x + gy.y * bw.y);
            });
            dr = block_sum<NT>(dr, red);
            dc = block_sum<NT>(dc, red);
        }
    }

    if constexpr (KIND == K_BATCH) {
        if (t == 0 && want_pos) {
            A.dpos[2 * (k - A.obase)] = (float)dr;
            A.dpos[2 * (k - A.obase) + 1] = (float)dc;
        }
    } else {
        // ---- sequential updates (STEP03_BACK.m:84-92, Run_MSET_recon.m:155-157)
        if (want_probe) {
            const BufC pr(A.probe, N2);
            const float s2 = A.s2;
            auto upd = staged<16>([&](const Pt& p) { return pr[p.e]; }, [&](const Pt& p, float2 v, float2 o) {
                const float2 w = cscale(v, inv_n2);
                pr.put(p.e, csub(o, cscale(w, s2)));
                return w;
            });
            auto unshift = [&](const Pt& p, float2 v) { return cmulc(v, phase(p.y, p.x)); };
            if (have_bw) {
                // MATLAB shifts back_wave_f: fftn(bw) .* conj(Phi) -> ifftn
                E.fwd(NoOp(), unshift);
                E.inv(NoOp(), upd);
            } else {
                E.inv(unshift, upd);
            }
        }
        if (t == 0 && want_pos) {
            A.xlist[k] = (float)(row + (double)A.s3 * dr);
            A.ylist[k] = (float)(col + (double)A.s3 * dc);
        }
        // the object step (every voxel of bin a moves by s1 * gradient plane a, STEP03_BACK.m:84) is the
        // host's next launch
    }
}

// ---------------------------------------------------------------------------------------------------
// Separable propagator, planes in global memory (the 256 x 256 engine).
//
// The Fresnel propagator of Func_generate_propagator.m, exp(-i pi lambda dz (qx^2 + qy^2)), is a product
// p(kx, ky) = px(kx) py(ky), so one propagation step is P = Px Py with Px = Fx^-1 px Fx acting along x
// alone and Py along y alone; they commute.  Alternating their order from plane to plane,
//     t0, Px Py, t1, Py Px, t2, Px Py, ...
// puts two half-steps along the same axis around every transmission multiply, so each chain pass runs
// [F, x p, F^-1, (store psi_a, x t_a), F, x p, F^-1] along one axis: a plane costs one pass over the
// global plane buffers instead of two.  The adjoint does the same with conj(p) and the gradient step.
// The 1D factors live in LDS, so the propagator costs no global loads.  The transmission is read as its
// phase r (in units of pi, reduced to [-1, 1]) and formed where it is applied, t = exp(i pi r): half the
// bytes of a complex plane.
//
// Orientation of pass X_k: columns (along y) for even k, rows (along x) for odd k.  The forward step of
// plane a is in X_(a+1), the adjoint step in X_(2L+2-a): the reference pairing conj(psi_(a+1)) reads each
// stored wave in the orientation that wrote it, while the transmission of plane a is read in opposite
// orientations, so it is kept in both layouts (tph: forward, tph_b: adjoint).  Gradient slab planes
// are in their adjoint pass's layout (odd planes transposed; the host puts them back after the sum).
// Arrays kept in the other layout than the pass that reads them (fftn(probe) and sqrt(M), [x][y], in a
// column pass; psi0, [y][x], written in a row pass) go through the LDS tile (TileLoad / TileStore).
//
// Checkpointed waves (WV = WAVES_CHECKPOINT): segments of K waves, s = (i - 1) / K for psi_i, i = 1 .. L.
// Pass X_(i+1) makes psi_i from its input I_(i+1); the forward sweep writes the input I_(sK+2) of each
// segment below the last straight into a checkpoint plane (the pass before it stores there instead of in
// the plane buffer) and keeps the last segment's waves.  On entering a lower segment the adjoint reruns its
// K forward passes from the checkpoint, their outputs alternating between two rebuild planes, each storing
// its wave: the forward sweep's own arithmetic, about one forward sweep more in all.
//
// Used for the batched gradient with the reference pairing and stored or checkpointed waves, and for the
// forward-only kernels; the sequential kernel, the exact pairing and the recomputed waves use
// process_position.
template <int N, int NT, int KIND, int WV = WAVES_STORE, class Eng>
__device__ __forceinline__ void process_sep(const PosArgs& A, int k, int slot, int slab, bool first, Eng& E,
                                            float2* phx, float2* phy, const float2* px, const float2* py,
                                            double* red) {
    static_assert(KIND == K_BATCH || KIND == K_FWD, "separable path: batched gradient or forward only");
    using CT = std::true_type;    // column pass
    using RT = std::false_type;   // row pass
    constexpr int N2 = N * N;
    constexpr int FWD = Eng::FWD, INV = Eng::INV, SWAP = Eng::ST_SWAP, NOST = Eng::ST_NONE;
    const int t = threadIdx.x;
    const int L = A.L;
    const float inv_n = 1.0f / (float)N, inv_n2 = 1.0f / (float)N2;
    constexpr bool GRAD = KIND == K_BATCH;
    constexpr bool ckpt = GRAD && WV == WAVES_CHECKPOINT;
    static_assert(WV == WAVES_STORE || WV == WAVES_CHECKPOINT, "separable path: stored or checkpointed waves");
    const int K = ckpt ? A.ckseg : 1, SL = ckpt ? (L - 1) / K : 0;
    // Per-slot scratch.  Store: plane i holds psi_i, i = 0 .. L.  Checkpoint: plane 0 psi0, planes 1 + s the
    // checkpoint inputs I_(sK+2) (s < SL), planes 1 + SL + (i - 1) % K the waves psi_i of the segment held,
    // planes 1 + SL + K and 2 + SL + K the rebuild passes' outputs.
    float2* save = A.save + (size_t)slot * (ckpt ? checkpoint_planes(L, K) : L + 2) * N2;
    auto wave = [&](int i) -> float2* {
        if constexpr (ckpt) return save + (size_t)(i > 0 ? 1 + SL + (i - 1) % K : 0) * N2;
        else return save + (size_t)i * N2;
    };
    // checkpoint: whether pass X_k's input I_k is a checkpoint, and where pass X_k's output I_(k+1) goes
    auto ck_in = [&](int kk) -> const float2* {
        return ckpt && kk >= 2 && (kk - 2) % K == 0 && (kk - 2) / K < SL ? save + (size_t)(1 + (kk - 2) / K) * N2 : nullptr;
    };
    auto ck_out = [&](int kk) -> float2* {
        return ckpt && (kk - 1) % K == 0 && (kk - 1) / K < SL ? save + (size_t)(1 + (kk - 1) / K) * N2 : nullptr;
    };
    // psi_i is kept by the forward sweep (checkpoint: only the last segment's)
    auto keeps = [&](int i) { return GRAD && (!ckpt || (i - 1) / K == SL); };
    E.bind(A.planes + (size_t)slot * 2 * N2);
    const bool want_pos = GRAD && (A.want & WANT_POS) != 0;
    const bool want_probe = GRAD && (A.want & WANT_PROBE) != 0;

    // ---- sub-pixel shift factors, separable (Func_generate_probe_wave_scan_sub_pixel.m:4-27)
    const double row = (double)A.xlist[k], col = (double)A.ylist[k];
    const double dxs = row - (double)A.cent, dys = col - (double)A.cent;
    for (int j = t; j < 2 * N; j += NT) {
        const int jj = j < N ? j : j - N;
        const double d = j < N ? dxs : dys;
        const double a = -2.0 * M_PI * d * shift_grid<N>(jj) / (double)N;
        double sn, cs;
        sincos(a, &sn, &cs);
        (j < N ? phx : phy)[jj] = make_float2((float)cs, (float)sn);
    }
    __syncthreads();
    auto phase = [&](int y, int x) { return cmul(phy[y], phx[x]); };
    // exp(i pi r), r in [-1, 1]: v_cos_f32 / v_sin_f32 take revolutions (r / 2 in [-0.5, 0.5]); their error
    // there is at most 1.3e-7 absolute (4M points against double), two float ulps at magnitude 1.  sincospif
    // (5e-8) costs more VALU time than the halved transmission bytes save.
    auto expi = [](float r) {
        const float h = 0.5f * r;
        return make_float2(__builtin_amdgcn_cosf(h), __builtin_amdgcn_sinf(h));
    };

    // propagation factors along the pass axis, with the 1/N of that axis' inverse transform
    auto pfac = [&](auto colt) {
        return [&](const Pt& p, float2 v) {
            return cscale(cmul(v, decltype(colt)::value ? py[p.y] : px[p.x]), inv_n);
        };
    };
    auto bfac = [&](auto colt) {
        return [&](const Pt& p, float2 v) {
            return cscale(cmulc(v, decltype(colt)::value ? py[p.y] : px[p.x]), inv_n);
        };
    };

    // ---- X0 (columns): fftn(probe) .* Phi, inverse along y.  fftn(probe) is [x][y]: through the LDS tile
    E.template chain<true, SWAP, INV>(tile_load(A.probeF, [&](const Pt& p, float2, float2 f) {
        return cmul(f, phase(p.y, p.x));
    }));
    // ---- X1 (rows): inverse along x -> psi0 (stored [y][x] through the LDS tile), x t_0, half-step along x
    {
        const BufF t0(A.tph, N2);
        auto scale = [&](const Pt&, float2 v) { return cscale(v, inv_n2); };
        auto times_t = staged<16>([&](const Pt& p) { return t0[p.e]; }, [&](const Pt&, float2 v, float o) { return cmul(v, expi(o)); });
        if constexpr (ckpt) E.template redirect<false>(nullptr, ck_out(1));
        E.template chain<false, SWAP, INV, FWD, INV>(NoOp(), seq(scale, TileStore{GRAD ? save : nullptr}, times_t),
                                                     pfac(RT{}));
    }
    // ---- X_(a+1), a = 1 .. L-1: finish plane a-1's step, store psi_a, x t_a, start plane a's step
    auto fwd_pass = [&](auto colt, int a) {
        constexpr bool COL = decltype(colt)::value;
        const BufF ta(A.tph + (size_t)a * N2, N2);
        const BufC sv(save + (size_t)a * N2, N2);
        auto mid = staged<16>([&](const Pt& p) { return ta[p.e]; }, [&](const Pt& p, float2 v, float o) {
            if (GRAD) sv.put(p.e, v);
            return cmul(v, expi(o));
        });
        E.template chain<COL, SWAP, FWD, INV, FWD, INV>(NoOp(), pfac(colt), mid, pfac(colt));
    };
    // the same pass with its input and output redirected and its wave kept or not (checkpoint variant)
    auto fwd_pass_io = [&](auto colt, int a, bool keep, const float2* in, float2* out) {
        constexpr bool COL = decltype(colt)::value;
        const BufF ta(A.tph + (size_t)a * N2, N2);
        const BufC sv(wave(a), N2);
        auto mid = staged<16>([&](const Pt& p) { return ta[p.e]; }, [&](const Pt& p, float2 v, float o) {
            if (keep) sv.put(p.e, v);
            return cmul(v, expi(o));
        });
        E.template redirect<COL>(in, out);
        E.template chain<COL, SWAP, FWD, INV, FWD, INV>(NoOp(), pfac(colt), mid, pfac(colt));
    };
    if constexpr (ckpt) {
        for (int a = 1; a < L; ++a) {
            if ((a + 1) % 2 == 0) fwd_pass_io(CT{}, a, keeps(a), ck_in(a + 1), ck_out(a + 1));
            else fwd_pass_io(RT{}, a, keeps(a), ck_in(a + 1), ck_out(a + 1));
        }
    } else {
        for (int a = 1; a < L; ++a) {
            if ((a + 1) % 2 == 0) fwd_pass(CT{}, a);
            else fwd_pass(RT{}, a);
        }
    }
    // ---- X_(L+1): finish plane L-1's step, store psi_L, forward transform along this axis
    auto last_pass = [&](auto colt) {
        constexpr bool COL = decltype(colt)::value;
        const BufC sv(wave(L), N2);
        auto keep = [&](const Pt& p, float2 v) {
            if (GRAD) sv.put(p.e, v);
            return v;
        };
        if constexpr (ckpt) E.template redirect<COL>(ck_in(L + 1), nullptr);
        E.template chain<COL, SWAP, FWD, INV, FWD>(NoOp(), pfac(colt), keep);
    };
    if ((L + 1) % 2 == 0) last_pass(CT{});
    else last_pass(RT{});
    if constexpr (ckpt) E.unbind();

    // checkpoint: rebuild segment s's waves psi_(sK+1) .. psi_(sK+K) by its forward passes X_(sK+2) ..
    // X_(sK+K+1) from the checkpoint I_(sK+2); the segment's last pass stops after its wave
    auto rebuild = [&](int s) {
        const float2* src = save + (size_t)(1 + s) * N2;
        for (int n = 0; n < K; ++n) {
            const int kk = s * K + 2 + n, i = kk - 1;
            if (n + 1 < K) {
                float2* dst = save + (size_t)(1 + SL + K + (n & 1)) * N2;
                if (kk % 2 == 0) fwd_pass_io(CT{}, i, true, src, dst);
                else fwd_pass_io(RT{}, i, true, src, dst);
                src = dst;
            } else {
                const BufC sv(wave(i), N2);
                auto keep = [&](const Pt& p, float2 v) {
                    sv.put(p.e, v);
                    return v;
                };
                if (kk % 2 == 0) {
                    E.template redirect<true>(src, nullptr);
                    E.template chain<true, NOST, FWD, INV>(NoOp(), pfac(CT{}), keep);
                } else {
                    E.template redirect<false>(src, nullptr);
                    E.template chain<false, NOST, FWD, INV>(NoOp(), pfac(RT{}), keep);
                }
            }
        }
        E.unbind();
    };

    // ---- X_(L+2): forward along the other axis -> exit wave Psi; intensity, error, residual
    // (STEP02_FORWARD_4D.m:36-49); the residual times conj(p) and the inverse along this axis start the adjoint
    double err = 0.0;
    {
        float* Iout = A.inten ? A.inten + (size_t)k * N2 : nullptr;
        const float* MT = A.sqrtM + (size_t)k * N2;                     // sqrt(M), [x][y]
        auto resid = [&](const Pt& p, float2 v, float m) {
            const float I = v.x * v.x + v.y * v.y;
            if (KIND == K_FWD && Iout) Iout[((p.y + N / 2) % N) * N + (p.x + N / 2) % N] = I;
            const float d = m - __builtin_amdgcn_sqrtf(I);
            err += (double)(d * d);
            if (KIND == K_FWD) return v;
            const float f = 1.0f - m * __builtin_amdgcn_rsqf(I + 1e-30f);
            return cmulc(make_float2(v.x * f, v.y * f), cmul(px[p.x], py[p.y]));
        };
        if (L % 2 == 0) {
            // column pass: sqrt(M) is [x][y], through the LDS tile
            auto ex = tile_load(MT, resid);
            if constexpr (GRAD) E.template chain<true, SWAP, FWD, INV>(NoOp(), ex);
            else E.template chain<true, NOST, FWD>(NoOp(), ex);
        } else {
            const BufF MB(MT, N2);
            auto ex = staged<16>([&](const Pt& p) { return MB[p.e]; }, resid);
            if constexpr (GRAD) E.template chain<false, SWAP, FWD, INV>(NoOp(), ex);
            else E.template chain<false, NOST, FWD>(NoOp(), ex);
        }
    }
    err = block_sum<NT>(err, red);
    if (t == 0) A.err_pos[k - A.obase] = err;
    if constexpr (KIND == K_FWD) return;

    // ---- adjoint (STEP03_BACK.m:12-24): plane a's step in X_(2L+2-a), columns for even a
    double dr = 0.0, dc = 0.0;
    auto adj_pass = [&](auto colt, int a) {
        constexpr bool COL = decltype(colt)::value;
        const BufF tb(A.tph_b + (size_t)a * N2, N2);
        const BufC ws(wave(a + 1), N2);                               // reference pairing: the wave after plane a
        const BufF g(A.gslab + ((size_t)slab * L + a) * N2, N2);
        const BufF gin(A.gslab + ((size_t)slab * L + a) * N2, first ? 0u : N2);   // empty view: reads 0
        auto mid = staged<8>([&](const Pt& p) { return AdjPh{tb[p.e], ws[p.e], gin[p.e]}; },
                              [&](const Pt& p, float2 u, const AdjPh& o) {
                                  const float2 tc = cconj(expi(o.ph));
                                  const float2 q = cmul(tc, cmulc(u, o.w));
                                  g.put(p.e, o.g - q.y);                // Re(1i * conj(t) * conj(psi) * b)
                                  return cmul(u, tc);
                              });
        auto scale = [&](const Pt&, float2 v) { return cscale(v, inv_n2); };
        const bool head = a == L - 1, tail = a == 0;
        if constexpr (COL) {
            // plane 0 is a column pass: the position terms (STEP03_BACK.m:35-38) ride on it, from bw =
            // conj(t_0) .* b and MATLAB gradient() of conj(psi0) (psi0 is [y][x]: neighbours are +-1, +-N)
            const float2* p0 = save;
            auto pos = [&](const Pt& p, float2 bw) {
                if (want_pos) {
                    const int y = p.y, x = p.x;
                    const int xp = min(x + 1, N - 1), xm = max(x - 1, 0), yp = min(y + 1, N - 1), ym = max(y - 1, 0);
                    const float2 gx = cscale(csub(p0[y * N + xp], p0[y * N + xm]), xp - xm == 2 ? 0.5f : 1.0f);
                    const float2 gy = cscale(csub(p0[yp * N + x], p0[ym * N + x]), yp - ym == 2 ? 0.5f : 1.0f);
                    dr += (double)(gx.x * bw.x + gx.y * bw.y);
                    dc += (double)(gy.x * bw.x + gy.y * bw.y);
                }
                return bw;
            };
            if (head && tail) E.template chain<true, SWAP, INV, FWD>(NoOp(), seq(scale, mid, pos));
            else if (head) E.template chain<true, SWAP, INV, FWD, INV>(NoOp(), seq(scale, mid), bfac(colt));
            else if (tail) E.template chain<true, SWAP, FWD, INV, FWD>(NoOp(), bfac(colt), seq(mid, pos));
            else E.template chain<true, SWAP, FWD, INV, FWD, INV>(NoOp(), bfac(colt), mid, bfac(colt));
        } else {
            if (head) E.template chain<false, SWAP, INV, FWD, INV>(NoOp(), seq(scale, mid), bfac(colt));
            else E.template chain<false, SWAP, FWD, INV, FWD, INV>(NoOp(), bfac(colt), mid, bfac(colt));
        }
    };
    if constexpr (ckpt) {
        int cseg = SL;     // the segment whose waves are held
        for (int a = L - 1; a >= 0; --a) {
            if (a / K != cseg) {
                cseg = a / K;                                         // psi_(a+1): segment (a + 1 - 1) / K
                rebuild(cseg);
            }
            if (a % 2 == 0) adj_pass(CT{}, a);
            else adj_pass(RT{}, a);
        }
    } else {
        for (int a = L - 1; a >= 0; --a) {
            if (a % 2 == 0) adj_pass(CT{}, a);
            else adj_pass(RT{}, a);
        }
    }
    if (want_pos) {
        dr = block_sum<NT>(dr, red);
        dc = block_sum<NT>(dc, red);
        if (t == 0) {
            A.dpos[2 * (k - A.obase)] = (float)dr;
            A.dpos[2 * (k - A.obase) + 1] = (float)dc;
        }
    }
    // ---- X_(2L+3) (rows): forward along x -> fftn(bw); the probe term in the Fourier domain,
    // shift(bw, -d) = ifftn(fftn(bw) .* conj(Phi)), added to the slab ([x][y])
    if (want_probe) {
        const BufC ps(A.pslab + (size_t)slab * N2, N2);
        const BufC psin(A.pslab + (size_t)slab * N2, first ? 0u : N2);   // empty view: reads 0
        auto term = staged<16>([&](const Pt& p) { return psin[p.e]; }, [&](const Pt& p, float2 v, float2 o) {
            ps.put(p.e, cadd(o, cmulc(v, phase(p.y, p.x))));
            return v;
        });
        E.template chain<false, NOST, FWD>(NoOp(), term);
    }
}

template <int N, int NT> __device__ __forceinline__ void load_twiddles(float2* tw, const float2* g) {
    for (int i = threadIdx.x; i < N; i += NT) tw[i] = g[i];
    __syncthreads();
}

#define MSET_LDS(N, NT, SH)                                      \
    using Eng = typename Cfg<N, SH>::Eng;                        \
    __shared__ float2 s[Eng::LDS_ELEMS];                         \
    __shared__ float2 tw[N], phx[N], phy[N];                     \
    __shared__ double red[(NT + 63) / 64];                       \
    load_twiddles<N, NT>(tw, A.tw);                              \
    Eng E;                                                       \
    E.init(s, tw);

// PFAC: the propagator as its 1D factors from LDS (a separable propagator), as in the batched kernel
template <int N, bool PFAC = false>
__global__ __launch_bounds__(Cfg<N>::NT, Cfg<N>::MINB) void k_plain(PosArgs A, int k) {
    constexpr int NT = Cfg<N>::NT;
    MSET_LDS(N, NT, 0)
    float2* pxs = nullptr;
    float2* pys = nullptr;
    if constexpr (PFAC) {
        __shared__ float2 fx[N], fy[N];
        for (int i = threadIdx.x; i < 2 * N; i += NT) (i < N ? fx : fy)[i % N] = A.p1d[i];
        __syncthreads();
        pxs = fx;
        pys = fy;
    }
    process_position<N, NT, K_PLAIN, WAVES_STORE, PFAC>(A, k, 0, 0, true, E, phx, phy, red, pxs, pys);
}

// WV: the exit-wave variant (A.waves), a separate instantiation each so that the store variant carries
// none of the others' code.  Position p of the launch adds to slab p % nslab.  The host launches either at most
// nslab positions or a grid that divides nslab, so every slab's positions run on one workgroup in
// increasing p: each slab sums the same positions in the same order whatever the grid size.
// SEP: the separable-propagator routine (process_sep; the 256 x 256 engine, stored waves), its 1D factors in LDS.
// SH: the workgroup shape (Cfg).  PFAC: the general routine applies the propagator as its 1D factors (LDS).
template <int N, int SH, int WV, bool SEP, bool PFAC = false>
__global__ __launch_bounds__((Cfg<N, SH>::NT), (Cfg<N, SH>::MINB)) void k_batch(PosArgs A) {
    constexpr int NT = Cfg<N, SH>::NT;
    MSET_LDS(N, NT, SH)
    if constexpr (SEP) {
        __shared__ float2 px[N], py[N];
        for (int i = threadIdx.x; i < 2 * N; i += NT) (i < N ? px : py)[i % N] = A.p1d[i];
        __syncthreads();
        for (int p = blockIdx.x; p < A.count; p += gridDim.x) {
            const int k = A.order ? A.order[A.first + p] : A.first + p;
            process_sep<N, NT, K_BATCH, WV>(A, k, blockIdx.x, p % A.nslab, p < A.nslab, E, phx, phy, px, py, red);
            __syncthreads();
        }
    } else {
        float2* pxs = nullptr;
        float2* pys = nullptr;
        if constexpr (PFAC) {
            __shared__ float2 fx[N], fy[N];
            for (int i = threadIdx.x; i < 2 * N; i += NT) (i < N ? fx : fy)[i % N] = A.p1d[i];
            __syncthreads();
            pxs = fx;
            pys = fy;
        }
        for (int p = blockIdx.x; p < A.count; p += gridDim.x) {
            const int k = A.order ? A.order[A.first + p] : A.first + p;
            const int slab = p % A.nslab;
            process_position<N, NT, K_BATCH, WV, PFAC>(A, k, blockIdx.x, slab, p < A.nslab, E, phx, phy, red, pxs, pys);
            __syncthreads();
        }
    }
}

// MT: the work items are positions of several tilts (A.tilts, A.order), so a whole error pass over a tilt
// series is one persistent launch; a separate instantiation, so the one-tilt kernel carries none of it
// PFAC (general routine below 256): the propagator as its 1D factors from LDS.
template <int N, bool MT, bool SEP, bool PFAC = false>
__global__ __launch_bounds__(Cfg<N>::NT, Cfg<N>::MINB) void k_fwd(PosArgs A) {
    constexpr int NT = Cfg<N>::NT;
    MSET_LDS(N, NT, 0)
    float2* px = nullptr;
    float2* py = nullptr;
    if constexpr (SEP || PFAC) {
        __shared__ float2 pxs[N], pys[N];
        for (int i = threadIdx.x; i < 2 * N; i += NT) (i < N ? pxs : pys)[i % N] = A.p1d[i];
        __syncthreads();
        px = pxs;
        py = pys;
    }
    for (int p = blockIdx.x; p < A.count; p += gridDim.x) {
        if constexpr (MT) {
            const int item = A.order[p];
            const TiltArgs& T = A.tilts[item >> ITEM_TILT_SHIFT];
            PosArgs B = A;
            B.trans = const_cast<float2*>(T.trans);
            B.tph = T.tph;
            B.probeF = T.probeF;
            B.sqrtM = T.sqrtM;
            B.xlist = T.xlist;
            B.ylist = T.ylist;
            B.err_pos = T.err_pos;
            B.obase = 0;
            if constexpr (SEP) process_sep<N, NT, K_FWD>(B, item & ITEM_POS_MASK, blockIdx.x, blockIdx.x, true, E, phx, phy, px, py, red);
            else process_position<N, NT, K_FWD, WAVES_STORE, PFAC>(B, item & ITEM_POS_MASK, blockIdx.x, blockIdx.x, true, E,
                                                                  phx, phy, red, px, py);
        } else {
            const int k = A.order ? A.order[A.first + p] : A.first + p;
            if constexpr (SEP) process_sep<N, NT, K_FWD>(A, k, blockIdx.x, blockIdx.x, true, E, phx, phy, px, py, red);
            else process_position<N, NT, K_FWD, WAVES_STORE, PFAC>(A, k, blockIdx.x, blockIdx.x, p == (int)blockIdx.x, E,
                                                                  phx, phy, red, px, py);
        }
        __syncthreads();
    }
}

// One plane:
//   dir 0: out = fftn(in), in row-major [y][x], out written transposed ([x][y]); alpha/beta ignored;
//   dir 1: out = beta * out + alpha * ifftn(in) (with 1/N^2), in transposed ([x][y]), out row-major.
template <int N>
__global__ __launch_bounds__(Cfg<N>::NT) void k_plane_fft(const float2* in, float2* out, const float2* twg, int dir,
                                                         float alpha, float beta, float2* planes, float2* out2,
                                                         float alpha2, float beta2) {
    constexpr int NT = Cfg<N>::NT;
    using Eng = typename Cfg<N>::Eng;
    __shared__ float2 s[Eng::LDS_ELEMS];
    __shared__ float2 tw[N];
    load_twiddles<N, NT>(tw, twg);
    Eng E;
    E.init(s, tw);
    E.bind(planes);
    if (dir == 0) {
        const BufC ib(in, N * N), ob(out, N * N);           // in [y][x], out [x][y]
        E.template fwd<true>([&](const Pt& p, float2) { return ib[p.e]; },
                             [&](const Pt& p, float2 v) {
                                 ob.put(p.e, v);
                                 return v;
                             });
    } else {
        const BufC ib(in, N * N), ob(out, N * N);           // in [x][y], out [y][x]
        const BufC ob2(out2, out2 ? N * N : 0);
        const float sc = alpha / (float)(N * N), sc2 = alpha2 / (float)(N * N);
        E.template inv<true>([&](const Pt& p, float2) { return ib[p.e]; },
                             [&](const Pt& p, float2 v) {
                                 const float2 w = cscale(v, sc);
                                 ob.put(p.e, beta == 0.f ? w : cadd(cscale(ob[p.e], beta), w));
                                 if (out2) {
                                     const float2 w2 = cscale(v, sc2);
                                     ob2.put(p.e, beta2 == 0.f ? w2 : cadd(cscale(ob2[p.e], beta2), w2));
                                 }
                                 return v;
                             });
    }
}

// One pass of the 2D transform of an N x N plane as batched line transforms, TL lines per workgroup:
// at N = 256 the plane does not fit in one workgroup's LDS, and a one-workgroup transform streams it
// through a single CU.  ROWS: transform the rows of `in` into `out` (same layout).  Otherwise transform
// the columns of `in` and store column j as row j of `out` (the transposed layouts of launch_plane_fft),
// with out = alpha/N^2 v + beta out for an inverse and the same into out2 with (alpha2, beta2).  Both
// passes keep line l at s[l (N + 1) + e]; loads and stores run along rows of the global arrays.
template <int N, int TL, int NT, bool INV, bool ROWS>
__global__ __launch_bounds__(NT) void k_plane_lines(const float2* __restrict__ in, float2* __restrict__ out,
                                                    const float2* __restrict__ twg, float alpha, float beta,
                                                    float2* __restrict__ out2, float alpha2, float beta2) {
    constexpr int LS = N + 1;
    __shared__ float2 s[LS * TL];
    __shared__ float2 tw[N];
    for (int i = threadIdx.x; i < N; i += NT) tw[i] = twg[i];
    const int l0 = blockIdx.x * TL;
    for (int i = threadIdx.x; i < N * TL; i += NT) {
        if (ROWS) {
            const int l = i / N, e = i % N;
            s[l * LS + e] = in[(size_t)(l0 + l) * N + e];
        } else {
            const int l = i % TL, e = i / TL;
            s[l * LS + e] = in[(size_t)e * N + l0 + l];
        }
    }
    __syncthreads();
    line_fft<N, NT, TL, 1, LS, 0, 1, INV>(s, tw);
    const float sc = alpha / (float)(N * N), sc2 = alpha2 / (float)(N * N);
    for (int i = threadIdx.x; i < N * TL; i += NT) {
        const int l = i / N, e = i % N;
        const float2 v = s[l * LS + e];
        const size_t g = (size_t)(l0 + l) * N + e;
        if (ROWS || !INV) {
            out[g] = v;
        } else {
            const float2 w = cscale(v, sc);
            out[g] = beta == 0.f ? w : cadd(cscale(out[g], beta), w);
            if (out2) {
                const float2 w2 = cscale(v, sc2);
                out2[g] = beta2 == 0.f ? w2 : cadd(cscale(out2[g], beta2), w2);
            }
        }
    }
}

// ------------------------------------------------------------------ host launchers
#define MSET_DISPATCH(N, BODY)                                              \
    switch (N) {                                                            \
        case 16: { constexpr int NN = 16; BODY; } break;                    \
        case 24: { constexpr int NN = 24; BODY; } break;                    \
        case 32: { constexpr int NN = 32; BODY; } break;                    \
        case 48: { constexpr int NN = 48; BODY; } break;                    \
        case 64: { constexpr int NN = 64; BODY; } break;                    \
        case 96: { constexpr int NN = 96; BODY; } break;                    \
        case 128: { constexpr int NN = 128; BODY; } break;                  \
        case 256: { constexpr int NN = 256; BODY; } break;                  \
        default: return hipErrorInvalidValue;                               \
    }

// the sizes below 256 (the kernels whose plane is on chip)
#define MSET_DISPATCH_SMALL(N, BODY)                                        \
    switch (N) {                                                            \
        case 16: { constexpr int NN = 16; BODY; } break;                    \
        case 24: { constexpr int NN = 24; BODY; } break;                    \
        case 32: { constexpr int NN = 32; BODY; } break;                    \
        case 48: { constexpr int NN = 48; BODY; } break;                    \
        case 64: { constexpr int NN = 64; BODY; } break;                    \
        case 96: { constexpr int NN = 96; BODY; } break;                    \
        case 128: { constexpr int NN = 128; BODY; } break;                  \
        default: return hipErrorInvalidValue;                               \
    }

}  // namespace mset
