// mset_oracle.cpp -- CPU restatement of the MSET multislice forward/adjoint path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it.  The product library (mset_amd/libmset_amd.so) never
// links, loads or calls it.
//
// It restates, step by step and in the MATLAB conventions (fftshift/ifftshift around every
// transform, 1-based centres round((N+1)/2), column-major data reversed into C order [z][y][x]):
//   Func_generate_propagator.m:17-38          -> oracle_propagator
//   Func_generate_probe_wave.m:14-56          -> oracle_probe
//   STEP00_INIT.m:214-218                     -> oracle_probe_normalise
//   Func_generate_probe_wave_scan_sub_pixel.m -> subpixel_shift
//   STEP02_FORWARD_4D.m:6-51                  -> forward
//   STEP03_BACK.m:6-39                        -> back
//   STEP02_FORWARD_4D_SSET.m:9-35             -> forward_sset, transmission_sset
//   STEP03_BACK_SSET.m:9-28                   -> back_sset
//   Run_MSET_recon.m:133-158 (device 0)       -> oracle_run_tilt   (sequential per-position updates)
//   Run_MSET_recon.m:207-222 (device 0)       -> oracle_error_tilt (forward + loss only)
//   Func_rot_3Dvol_FourierShear.m, Func_inv_rot_3Dvol_FourierShear.m,
//   Func_rotate_3Dvol_ROTmat_FourierShear.m:14-133, SheerTransforms_for3Drot.m:16-49,
//   calculate_S_matrices_for3DFourierShearRot_perm{,_backward}_ver1.m, MatrixQuaternionRot.m,
//   My_obtain_quaternion_from_matrix.m        -> oracle_rotate
//   TV_method.m:5-84                           -> tv_method
// plus an "exact" gradient mode (the true adjoint of the forward model; the reference's
// STEP03_BACK.m:14-15 pairs the back-wave of plane a with the wave *after* plane a, which is not
// the gradient of its own loss) and the gradient-at-fixed-state sums used by the batched path.
//
// How each is pinned (tests/test_oracle.py): the propagator, probe, sub-pixel shift, transmission,
// forward waves, losses and rotation against the reference's CuAu example vectors (tests/golden/);
// the exact gradients against central finite differences; the reference-mode object term by the
// pairing identity with the exact one; the reference-mode position term against numpy.gradient (the
// MATLAB edge rule) on a wave with energy at the field border; oracle_run_tilt against a loop of
// one-position gradients that advances the state after each position; the SSET pair by reduction to
// the multislice one and finite differences; tv_method by the ROF closed form and its invariants; the
// rotation also by 90-degree permutations and round trips.  No function is left unpinned.
//
// Every function is templated on the real type: double is the correctness oracle, float restates
// the reference's single-precision arithmetic and is the timed CPU baseline.  FFTs are a
// self-contained mixed-radix Stockham FFT with long-double twiddles (no FFTW here).
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <typename T> using cx = std::complex<T>;

// ------------------------------------------------------------------------------------------------
// FFT
// ------------------------------------------------------------------------------------------------
template <typename T> struct Plan {
    int n = 0;
    std::vector<int> fac;
    std::vector<cx<T>> wf, wb;   // exp(-2 pi i k/n), exp(+2 pi i k/n)
};

template <typename T> const Plan<T>& get_plan(int n) {
    thread_local const Plan<T>* last = nullptr;
    if (last && last->n == n) return *last;
    static std::mutex mu;
    static std::map<int, Plan<T>*> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(n);
    if (it != cache.end()) { last = it->second; return *last; }
    auto* p = new Plan<T>();
    p->n = n;
    int m = n;
    for (int r : {4, 2, 3, 5}) while (m % r == 0) { p->fac.push_back(r); m /= r; }
    for (int r = 7; m > 1; r += 2) while (m % r == 0) { p->fac.push_back(r); m /= r; }
    p->wf.resize(n); p->wb.resize(n);
    for (int k = 0; k < n; ++k) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        p->wf[k] = cx<T>((T)cosl(a), (T)-sinl(a));
        p->wb[k] = cx<T>((T)cosl(a), (T)sinl(a));
    }
    cache[n] = p;
    last = p;
    return *p;
}

// Iterative Stockham autosort, radix sequence P.fac, twiddles from the size-n table w.
// Stage with radix R and span Ns: butterfly j reads in[j + r n/R], multiplies by w^(k r n/(Ns R)),
// k = j mod Ns, applies the R-point DFT and writes out[(j/Ns) Ns R + k + r Ns].
template <typename T> inline void dft_small(cx<T>* v, int R, const cx<T>* w, int n, cx<T>* t) {
    if (R == 2) {
        const cx<T> a = v[0], b = v[1];
        v[0] = a + b; v[1] = a - b;
    } else if (R == 4) {
        const cx<T> j = w[n / 4];   // -i forward, +i inverse
        const cx<T> a0 = v[0] + v[2], a1 = v[0] - v[2], b0 = v[1] + v[3], b1 = (v[1] - v[3]) * j;
        v[0] = a0 + b0; v[1] = a1 + b1; v[2] = a0 - b0; v[3] = a1 - b1;
    } else if (R == 3) {
        const cx<T> w1 = w[n / 3];
        const T c = w1.real(), sn = w1.imag();   // cos(2pi/3) = -1/2, +-sin(2pi/3)
        const cx<T> s = v[1] + v[2], d = v[1] - v[2];
        const cx<T> m = v[0] + c * s;
        const cx<T> r = cx<T>(-sn * d.imag(), sn * d.real());
        v[0] = v[0] + s; v[1] = m + r; v[2] = m - r;
    } else {
        for (int q = 0; q < R; ++q) {
            cx<T> acc(0, 0);
            for (int r = 0; r < R; ++r) acc += v[r] * w[(size_t)((q * r) % R) * (n / R)];
            t[q] = acc;
        }
        for (int q = 0; q < R; ++q) v[q] = t[q];
    }
}

template <typename T> void stockham(cx<T>* x, cx<T>* y, int n, const Plan<T>& P, const cx<T>* w) {
    cx<T>* in = x;
    cx<T>* out = y;
    int Ns = 1;
    int Rmax = 1;
    for (int R : P.fac) Rmax = std::max(Rmax, R);
    cx<T> vs[16], ts[16];
    std::vector<cx<T>> vh, th;
    cx<T>* v = vs;
    cx<T>* t = ts;
    if (Rmax > 16) { vh.resize(Rmax); th.resize(Rmax); v = vh.data(); t = th.data(); }
    for (int R : P.fac) {
        const int nr = n / R;
        const int tws = n / (Ns * R);
        for (int j = 0; j < nr; ++j) {
            const int k = j % Ns;
            for (int r = 0; r < R; ++r) v[r] = in[j + r * nr];
            if (k) for (int r = 1; r < R; ++r) v[r] *= w[(size_t)k * r * tws];
            dft_small<T>(v, R, w, n, t);
            const int base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; ++r) out[base + r * Ns] = v[r];
        }
        std::swap(in, out);
        Ns *= R;
    }
    if (in != x) std::memcpy(x, in, sizeof(cx<T>) * n);
}

// In-place 1D transform of n elements at stride `st`; sign -1 forward, +1 inverse (scaled 1/n).
template <typename T> void fft1(cx<T>* x, int n, int st, int sign, cx<T>* buf_in, cx<T>* buf_out) {
    const Plan<T>& P = get_plan<T>(n);
    for (int i = 0; i < n; ++i) buf_in[i] = x[(size_t)i * st];
    stockham<T>(buf_in, buf_out, n, P, sign < 0 ? P.wf.data() : P.wb.data());
    if (sign > 0) {
        const T s = T(1) / T(n);
        for (int i = 0; i < n; ++i) x[(size_t)i * st] = buf_in[i] * s;
    } else {
        for (int i = 0; i < n; ++i) x[(size_t)i * st] = buf_in[i];
    }
}

// Transform along one axis of a C-order [n2][n1][n0] array (axis 0 = fastest).
template <typename T> void fft_axis(cx<T>* a, int n0, int n1, int n2, int axis, int sign) {
    const int n = axis == 0 ? n0 : axis == 1 ? n1 : n2;
    const size_t st = axis == 0 ? 1 : axis == 1 ? (size_t)n0 : (size_t)n0 * n1;
    const int lines = n0 * n1 * n2 / n;
    get_plan<T>(n);
    // small transforms run on the calling thread: thread start-up would dominate them
#pragma omp parallel if ((long long)lines * n >= 4096)
    {
        std::vector<cx<T>> bi(n), bo(n);
#pragma omp for schedule(static)
        for (int l = 0; l < lines; ++l) {
            size_t base;
            if (axis == 0) base = (size_t)l * n0;
            else if (axis == 1) base = (size_t)(l / n0) * n0 * n1 + (l % n0);
            else base = (size_t)l;
            fft1<T>(a + base, n, (int)st, sign, bi.data(), bo.data());
        }
    }
}

// MATLAB fftshift / ifftshift of a 2D [ny][nx] plane (circshift by floor(n/2), resp. -floor(n/2)).
template <typename T> void shift2(cx<T>* a, int ny, int nx, bool inverse) {
    thread_local std::vector<cx<T>> t;
    t.assign(a, a + (size_t)nx * ny);
    const int sy = inverse ? (ny - ny / 2) % ny : ny / 2, sx = inverse ? (nx - nx / 2) % nx : nx / 2;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) a[(size_t)((y + sy) % ny) * nx + (x + sx) % nx] = t[(size_t)y * nx + x];
}

template <typename T> void shift3(cx<T>* a, int nx, int ny, int nz, bool inverse) {
    std::vector<cx<T>> t(a, a + (size_t)nx * ny * nz);
    const int sz = inverse ? (nz - nz / 2) % nz : nz / 2;
    const int sy = inverse ? (ny - ny / 2) % ny : ny / 2, sx = inverse ? (nx - nx / 2) % nx : nx / 2;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                a[((size_t)((z + sz) % nz) * ny + (y + sy) % ny) * nx + (x + sx) % nx] = t[((size_t)z * ny + y) * nx + x];
}

template <typename T> void fft2(cx<T>* a, int n, int sign) {
    fft_axis<T>(a, n, n, 1, 0, sign);
    fft_axis<T>(a, n, n, 1, 1, sign);
}

// fftshift(fftn(ifftshift(a))) (sign -1) or fftshift(ifftn(ifftshift(a))) (sign +1)
template <typename T> void fft2c(cx<T>* a, int n, int sign) {
    shift2<T>(a, n, n, true);
    fft2<T>(a, n, sign);
    shift2<T>(a, n, n, false);
}

inline int mround(double v) { return (int)std::floor(v + 0.5); }   // MATLAB round for v >= 0

// ------------------------------------------------------------------------------------------------
// Operators
// ------------------------------------------------------------------------------------------------
struct Geo {
    int n;          // lateral side (probe side == volume lateral side)
    int nz;         // volume depth
    int bin;        // slice_binning
    int L;          // numPlanes = ceil(nz / bin)
    double sigma;   // interaction constant, rad/(V A)
    double dx;      // potential_pixelsize, A
    bool sset;      // single-slice operator: one plane holding the whole depth, no propagation
};

inline int bin_lo(const Geo& g, int a) { return a * g.bin; }
inline int bin_hi(const Geo& g, int a) { return std::min((a + 1) * g.bin, g.nz); }

// Func_generate_probe_wave_scan_sub_pixel.m:4-27 : ifftn(fftn(p) .* ifftshift(exp(-2 pi i (dx X/nx + dy Y/ny))))
// with [Y, X] = meshgrid(-nx/2:nx/2-1, -ny/2:ny/2-1), X along dim 1 (our x), dx = x1 - round((n+1)/2).
template <typename T> void subpixel_shift(const cx<T>* p, cx<T>* out, int n, double x1, double y1) {
    const double cent = mround((n + 1) / 2.0);
    const T dx = (T)(x1 - cent), dy = (T)(y1 - cent);
    std::vector<cx<T>> F(p, p + (size_t)n * n);
    fft2<T>(F.data(), n, -1);
    // centred-grid value at MATLAB index i (1-based): -n/2 + (i-1); ifftshift: unshifted j <- centred (j + n/2) mod n
    for (int jy = 0; jy < n; ++jy) {
        const T gy = (T)(-n / 2.0 + (double)((jy + n / 2) % n));
        for (int jx = 0; jx < n; ++jx) {
            const T gx = (T)(-n / 2.0 + (double)((jx + n / 2) % n));
            const T arg = dx * gx / (T)n + dy * gy / (T)n;
            const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)arg;
            F[(size_t)jy * n + jx] *= cx<T>((T)cosl(ang), (T)sinl(ang));
        }
    }
    fft2<T>(F.data(), n, +1);
    std::memcpy(out, F.data(), sizeof(cx<T>) * n * n);
}

// STEP02_FORWARD_4D.m:12-20 : trans(:,:,a) = exp(1i*sigma*sum(V(bin a))*dx)
template <typename T> void transmission(const Geo& g, const T* V, cx<T>* trans) {
    const size_t n2 = (size_t)g.n * g.n;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n2; ++i) {
        for (int a = 0; a < g.L; ++a) {
            T s = 0;
            for (int z = bin_lo(g, a); z < bin_hi(g, a); ++z) s += V[(size_t)z * n2 + i];
            const T potz = s * (T)g.dx;
            const T ph = (T)g.sigma * potz;
            trans[(size_t)a * n2 + i] = cx<T>(std::cos(ph), std::sin(ph));
        }
    }
}

// STEP02_FORWARD_4D.m:31-49.  psi0 = incident wave (already shifted), prop = MATLAB-layout propagator,
// M = measured pattern (DC centred).  save: L planes of the wave after each plane; exit = centred Psi.
template <typename T>
double forward(const Geo& g, const cx<T>* trans, const cx<T>* psi0, const cx<T>* prop, const T* M,
               cx<T>* save, cx<T>* exitw, cx<T>* resi) {
    const size_t n2 = (size_t)g.n * g.n;
    std::vector<cx<T>> w(psi0, psi0 + n2);
    for (int a = 0; a < g.L; ++a) {
        for (size_t i = 0; i < n2; ++i) w[i] *= trans[(size_t)a * n2 + i];
        fft2c<T>(w.data(), g.n, -1);
        for (size_t i = 0; i < n2; ++i) w[i] *= prop[i];
        fft2c<T>(w.data(), g.n, +1);
        std::memcpy(save + (size_t)a * n2, w.data(), sizeof(cx<T>) * n2);
    }
    fft2c<T>(w.data(), g.n, -1);
    double err = 0.0;
    for (size_t i = 0; i < n2; ++i) {
        const T I = std::norm(w[i]);
        const T sm = std::sqrt(M[i]);
        const T si = std::sqrt(I);
        if (resi) resi[i] = w[i] * (T(1) - sm / std::sqrt(I + (T)1e-30));
        const T d = sm - si;
        err += (double)(d * d);
        if (exitw) exitw[i] = w[i];
    }
    return err / (double)n2;   // mean(abs(tmp_error(:)))
}

// MATLAB gradient(): central difference in the interior, one-sided at the edges, unit spacing.
template <typename T> void matlab_gradient(const cx<T>* f, int n, cx<T>* gx, cx<T>* gy) {
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const size_t i = (size_t)y * n + x;
            if (x == 0) gx[i] = f[i + 1] - f[i];
            else if (x == n - 1) gx[i] = f[i] - f[i - 1];
            else gx[i] = (f[i + 1] - f[i - 1]) / T(2);
            if (y == 0) gy[i] = f[i + n] - f[i];
            else if (y == n - 1) gy[i] = f[i] - f[i - n];
            else gy[i] = (f[i + n] - f[i - n]) / T(2);
        }
}

enum { MODE_REF = 0, MODE_EXACT = 1, MODE_SSET = 2 };

// Position update terms (STEP03_BACK.m:35-38): (drow, dcol) = Re sum(d_x conj(psi0) .* bw), Re sum(d_y
// conj(psi0) .* bw); d is MATLAB gradient() (mode REF) or the spectral derivative of the shift (EXACT).
template <typename T>
void position_terms(const Geo& g, int mode, const cx<T>* psi0, const cx<T>* bw, double* drow, double* dcol) {
    const size_t n2 = (size_t)g.n * g.n;
    const int n = g.n;
    std::vector<cx<T>> cp(n2), dX(n2), dY(n2);
    if (mode == MODE_REF) {
        for (size_t i = 0; i < n2; ++i) cp[i] = std::conj(psi0[i]);
        matlab_gradient<T>(cp.data(), n, dX.data(), dY.data());
    } else {
        // d psi0 / d(row) = ifftn(fftn(p) Phi (-2 pi i gx / n)) and psi0 = ifftn(fftn(p) Phi) so
        // d psi0/d(row) = ifftn(fftn(psi0) (-2 pi i gx/n)).  The update term uses -conj(d psi0/d row).
        std::vector<cx<T>> F(psi0, psi0 + n2);
        fft2<T>(F.data(), n, -1);
        std::vector<cx<T>> Fx(n2), Fy(n2);
        for (int jy = 0; jy < n; ++jy) {
            const double gy = -n / 2.0 + (double)((jy + n / 2) % n);
            for (int jx = 0; jx < n; ++jx) {
                const double gx = -n / 2.0 + (double)((jx + n / 2) % n);
                const size_t i = (size_t)jy * n + jx;
                const double c = 2.0 * M_PI / n;
                Fx[i] = F[i] * cx<T>(0, (T)(-c * gx));
                Fy[i] = F[i] * cx<T>(0, (T)(-c * gy));
            }
        }
        fft2<T>(Fx.data(), n, +1);
        fft2<T>(Fy.data(), n, +1);
        for (size_t i = 0; i < n2; ++i) { dX[i] = -std::conj(Fx[i]); dY[i] = -std::conj(Fy[i]); }
    }
    double sx = 0, sy = 0;
    T fx = 0, fy = 0;
    for (size_t i = 0; i < n2; ++i) {
        if (std::is_same<T, float>::value) { fx += std::real(dX[i] * bw[i]); fy += std::real(dY[i] * bw[i]); }
        else { sx += (double)std::real(dX[i] * bw[i]); sy += (double)std::real(dY[i] * bw[i]); }
    }
    if (std::is_same<T, float>::value) { sx = fx; sy = fy; }
    *drow = sx;
    *dcol = sy;
}


// STEP03_BACK.m:9-39.  Writes the complex gradient grad (n2*nz, replicated over each bin) and the final
// back-propagated wave bw = back_wave_f (real space, MATLAB layout).  Returns the position update terms
// (drow, dcol) = Re sum(d_x conj(psi0) .* bw), Re sum(d_y conj(psi0) .* bw) without the step factor.
// mode REF uses save(a) (STEP03_BACK.m:14-15); EXACT uses the wave entering plane a and the spectral
// derivative of the Fourier-shift for the position terms.
template <typename T>
void back(const Geo& g, int mode, const cx<T>* trans, const cx<T>* save, const cx<T>* psi0,
          const cx<T>* resi, const cx<T>* back_prop, cx<T>* grad, cx<T>* bw, double x1, double y1,
          double* drow, double* dcol) {
    const size_t n2 = (size_t)g.n * g.n;
    std::vector<cx<T>> b(resi, resi + n2);
    for (int a = g.L - 1; a >= 0; --a) {
        for (size_t i = 0; i < n2; ++i) b[i] *= back_prop[i];
        fft2c<T>(b.data(), g.n, +1);
        const cx<T>* ws = (mode == MODE_REF) ? save + (size_t)a * n2 : (a == 0 ? psi0 : save + (size_t)(a - 1) * n2);
        cx<T>* g0 = grad + (size_t)bin_lo(g, a) * n2;
        for (size_t i = 0; i < n2; ++i)
            g0[i] = cx<T>(0, 1) * std::conj(trans[(size_t)a * n2 + i]) * std::conj(ws[i]) * b[i];
        for (int z = bin_lo(g, a) + 1; z < bin_hi(g, a); ++z) std::memcpy(grad + (size_t)z * n2, g0, sizeof(cx<T>) * n2);
        for (size_t i = 0; i < n2; ++i) b[i] *= std::conj(trans[(size_t)a * n2 + i]);
        fft2c<T>(b.data(), g.n, -1);
    }
    fft2c<T>(b.data(), g.n, +1);
    std::memcpy(bw, b.data(), sizeof(cx<T>) * n2);

    position_terms<T>(g, mode, psi0, bw, drow, dcol);
    (void)x1; (void)y1;
}

// STEP03_BACK.m:84-86 : probe - step2 * shift(bw, 2*round((n+1)/2) - [row col])
template <typename T>
void probe_update(cx<T>* probe, const cx<T>* bw, int n, double x1, double y1, double s2) {
    const double c2 = 2.0 * mround((n + 1) / 2.0);
    std::vector<cx<T>> sh((size_t)n * n);
    subpixel_shift<T>(bw, sh.data(), n, (double)(T)(c2 - x1), (double)(T)(c2 - y1));
    for (size_t i = 0; i < (size_t)n * n; ++i) probe[i] -= (T)s2 * sh[i];
}

// STEP02_FORWARD_4D_SSET.m:9-35 : the projected potential of the whole depth as one transmission, no
// propagation: Psi = fftn(psi0 .* proj_trans); residual and error as STEP02_FORWARD_4D.m.
template <typename T>
double forward_sset(const Geo& g, const cx<T>* ptrans, const cx<T>* psi0, const T* M, cx<T>* resi) {
    const size_t n2 = (size_t)g.n * g.n;
    std::vector<cx<T>> w(n2);
    for (size_t i = 0; i < n2; ++i) w[i] = psi0[i] * ptrans[i];
    fft2c<T>(w.data(), g.n, -1);
    double err = 0.0;
    for (size_t i = 0; i < n2; ++i) {
        const T I = std::norm(w[i]);
        const T sm = std::sqrt(M[i]);
        if (resi) resi[i] = w[i] * (T(1) - sm / std::sqrt(I + (T)1e-30));
        const T d = sm - std::sqrt(I);
        err += (double)(d * d);
    }
    return err / (double)n2;
}

// STEP03_BACK_SSET.m:9-28 : b = ifftn(resi); grad = 1i conj(proj_trans) conj(psi0) b on every slice;
// back_wave_f = b .* conj(proj_trans); position terms as STEP03_BACK.m.
template <typename T>
void back_sset(const Geo& g, int mode, const cx<T>* ptrans, const cx<T>* psi0, const cx<T>* resi, cx<T>* grad,
               cx<T>* bw, double* drow, double* dcol) {
    const size_t n2 = (size_t)g.n * g.n;
    std::vector<cx<T>> b(resi, resi + n2);
    fft2c<T>(b.data(), g.n, +1);
    for (size_t i = 0; i < n2; ++i) grad[i] = cx<T>(0, 1) * std::conj(ptrans[i]) * std::conj(psi0[i]) * b[i];
    for (int z = 1; z < g.nz; ++z) std::memcpy(grad + (size_t)z * n2, grad, sizeof(cx<T>) * n2);
    for (size_t i = 0; i < n2; ++i) bw[i] = b[i] * std::conj(ptrans[i]);
    position_terms<T>(g, mode, psi0, bw, drow, dcol);
}

// STEP02_FORWARD_4D_SSET.m:9 : proj_trans = exp(1i*sigma*sum(RVol,3)*dx), evaluated left to right
template <typename T> void transmission_sset(const Geo& g, const T* V, cx<T>* ptrans) {
    const size_t n2 = (size_t)g.n * g.n;
    for (size_t i = 0; i < n2; ++i) {
        T s = 0;
        for (int z = 0; z < g.nz; ++z) s += V[(size_t)z * n2 + i];
        const T ph = ((T)g.sigma * s) * (T)g.dx;
        ptrans[i] = cx<T>(std::cos(ph), std::sin(ph));
    }
}

// the operator pair of the selected method
template <typename T> void trans_any(const Geo& g, const T* V, cx<T>* trans) {
    if (g.sset) transmission_sset<T>(g, V, trans);
    else transmission<T>(g, V, trans);
}
template <typename T>
double forward_any(const Geo& g, const cx<T>* trans, const cx<T>* psi0, const cx<T>* prop, const T* M, cx<T>* save,
                   cx<T>* resi) {
    if (g.sset) return forward_sset<T>(g, trans, psi0, M, resi);
    return forward<T>(g, trans, psi0, prop, M, save, nullptr, resi);
}
template <typename T>
void back_any(const Geo& g, int mode, const cx<T>* trans, const cx<T>* save, const cx<T>* psi0, const cx<T>* resi,
              const cx<T>* back_prop, cx<T>* grad, cx<T>* bw, double row, double col, double* dr, double* dc) {
    if (g.sset) back_sset<T>(g, mode, trans, psi0, resi, grad, bw, dr, dc);
    else back<T>(g, mode, trans, save, psi0, resi, back_prop, grad, bw, row, col, dr, dc);
}

// ------------------------------------------------------------------------------------------------
// Rotation (Fourier shear)
// ------------------------------------------------------------------------------------------------
// MatrixQuaternionRot.m:1-11 ; matrices are row-major here, M[r*3+c] = MAT(r+1, c+1)
void quat_rot(const double v_in[3], double theta_deg, double M[9]) {
    const double th = theta_deg * M_PI / 180.0;
    const double nv = std::sqrt(v_in[0] * v_in[0] + v_in[1] * v_in[1] + v_in[2] * v_in[2]);
    const double v0 = v_in[0] / nv, v1 = v_in[1] / nv, v2 = v_in[2] / nv;
    const double w = std::cos(th / 2), x = -std::sin(th / 2) * v0, y = -std::sin(th / 2) * v1, z = -std::sin(th / 2) * v2;
    const double R[9] = {1 - 2 * y * y - 2 * z * z, 2 * x * y + 2 * w * z, 2 * x * z - 2 * w * y,
                         2 * x * y - 2 * w * z, 1 - 2 * x * x - 2 * z * z, 2 * y * z + 2 * w * x,
                         2 * x * z + 2 * w * y, 2 * y * z - 2 * w * x, 1 - 2 * x * x - 2 * y * y};
    std::memcpy(M, R, sizeof(R));
}

void matmul3(const double A[9], const double B[9], double C[9]) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += A[r * 3 + k] * B[k * 3 + c];
            C[r * 3 + c] = s;
        }
}

// My_obtain_quaternion_from_matrix.m:11-39
void quat_from_matrix(const double M[9], double& qx, double& qy, double& qz, double& qw) {
    auto m = [&](int r, int c) { return M[(r - 1) * 3 + (c - 1)]; };
    const double Tr = m(1, 1) + m(2, 2) + m(3, 3);
    if (Tr > 0) {
        const double S = std::sqrt(Tr + 1) * 2;
        qw = 0.25 * S; qx = (m(3, 2) - m(2, 3)) / S; qy = (m(1, 3) - m(3, 1)) / S; qz = (m(2, 1) - m(1, 2)) / S;
    } else if (m(1, 1) > m(2, 2) && m(1, 1) > m(3, 3)) {
        const double S = std::sqrt(1.0 + m(1, 1) - m(2, 2) - m(3, 3)) * 2;
        qw = (m(3, 2) - m(2, 3)) / S; qx = 0.25 * S; qy = (m(1, 2) + m(2, 1)) / S; qz = (m(1, 3) + m(3, 1)) / S;
    } else if (m(2, 2) > m(3, 3)) {
        const double S = std::sqrt(1.0 + m(2, 2) - m(1, 1) - m(3, 3)) * 2;
        qw = (m(1, 3) - m(3, 1)) / S; qx = (m(1, 2) + m(2, 1)) / S; qy = 0.25 * S; qz = (m(2, 3) + m(3, 2)) / S;
    } else {
        const double S = std::sqrt(1.0 + m(3, 3) - m(1, 1) - m(2, 2)) * 2;
        qw = (m(2, 1) - m(1, 2)) / S; qx = (m(1, 3) + m(3, 1)) / S; qy = (m(2, 3) + m(3, 2)) / S; qz = 0.25 * S;
    }
}

// calculate_S_matrices_for3DFourierShearRot_perm_ver1.m:16-92 (backward=false) and
// ..._backward_ver1.m:16-115 (backward=true).  Returns absSum (inf if singular), coefficients a..h.
double shear_coeffs(double Xi, double Yi, double Zi, double W, int numPerm, bool backward, double c[8]) {
    double P3[3] = {Xi, Yi, Zi}, Q[3];
    for (int i = 0; i < 3; ++i) Q[i] = P3[(i + numPerm) % 3];   // circshift(.., [0 -numPerm])
    const double X = Q[0], Y = Q[1], Z = Q[2];
    if (backward) W = -W;
    double a, b, cc, d, e, f, g, h;
    if (std::fabs(Y * Z - X * W) > 1e-5 && std::fabs(X * Y - Z * W) > 1e-5) {
        a = (X * X + Y * Y) / (Y * Z - X * W);
        b = (Y * Y - Z * Z) / (X * Y - Z * W) - 2 * (Y * Z * (X * X + Y * Y)) / (X * Y - Z * W) / (Y * Z - X * W);
        cc = 2 * Y * Z / (X * Y - Z * W);
        d = 2 * (X * W - Y * Z);
        e = 2 * (X * Y - Z * W);
        f = -2 * X * Y / (Y * Z - X * W);
        g = (X * X - Y * Y) / (Y * Z - X * W) + 2 * X * Y * (Y * Y + Z * Z) / (X * Y - Z * W) / (Y * Z - X * W);
        h = (-1 + X * X + W * W) / (X * Y - Z * W);
    } else if (std::fabs(Y) < 1e-6) {
        a = -X / W; b = Z / W; cc = 0; d = 2 * X * W; e = -2 * Z * W; f = 0; g = -X / W; h = Z / W;
    } else {
        return INFINITY;
    }
    double v[8] = {a, b, cc, d, e, f, g, h};
    double s = 0;
    for (int i = 0; i < 8; ++i) { c[i] = backward ? -v[i] : v[i]; s += std::fabs(v[i]); }
    return s;
}

struct ShearPlan {
    int numStep = 0;
    int dims[4] = {0, 0, 0, 0};       // shear dimension (1..3) of the 4 shears in application order
    double ab[4][2] = {{0}};          // (a, b) coefficient pair of each shear
};

// Func_rotate_3Dvol_ROTmat_FourierShear.m:28-106
ShearPlan plan_rotation(const double MAT[9], double rotStepInit) {
    double xo, yo, zo, wo;
    quat_from_matrix(MAT, xo, yo, zo, wo);
    const double Ang = std::acos(wo) * 180.0 / M_PI * 2;
    const double nv = std::sqrt(xo * xo + yo * yo + zo * zo);
    const double V[3] = {xo / nv, yo / nv, zo / nv};
    ShearPlan P;
    P.numStep = (int)std::floor(Ang / rotStepInit) + 1;
    const double rotStep = Ang / P.numStep;
    const double w = std::cos(rotStep / 2 * M_PI / 180), s = std::sin(rotStep / 2 * M_PI / 180);
    const double x = -s * V[0], y = -s * V[1], z = -s * V[2];
    double af[3], ab_[3], tmp[8];
    for (int i = 1; i <= 3; ++i) af[i - 1] = shear_coeffs(x, y, z, w, i, false, tmp);
    for (int i = 1; i <= 3; ++i) ab_[i - 1] = shear_coeffs(x, y, z, w, i, true, tmp);
    int iF = 0, iB = 0;
    for (int i = 1; i < 3; ++i) { if (af[i] < af[iF]) iF = i; if (ab_[i] < ab_[iB]) iB = i; }
    const bool fwd = af[iF] < ab_[iB];
    const int mI = (fwd ? iF : iB) + 1;
    double C[8];
    shear_coeffs(x, y, z, w, mI, !fwd, C);
    const double A = C[0], B = C[1], Cc = C[2], D = C[3], E = C[4], F = C[5], G = C[6], H = C[7];
    if (fwd) {
        P.dims[0] = (mI + 1) % 3 + 1; P.ab[0][0] = G; P.ab[0][1] = H;
        P.dims[1] = mI % 3 + 1;       P.ab[1][0] = E; P.ab[1][1] = F;
        P.dims[2] = (mI + 2) % 3 + 1; P.ab[2][0] = Cc; P.ab[2][1] = D;
        P.dims[3] = (mI + 1) % 3 + 1; P.ab[3][0] = A; P.ab[3][1] = B;
    } else {
        P.dims[0] = (mI + 1) % 3 + 1; P.ab[0][0] = A; P.ab[0][1] = B;
        P.dims[1] = (mI + 2) % 3 + 1; P.ab[1][0] = Cc; P.ab[1][1] = D;
        P.dims[2] = mI % 3 + 1;       P.ab[2][0] = E; P.ab[2][1] = F;
        P.dims[3] = (mI + 1) % 3 + 1; P.ab[3][0] = G; P.ab[3][1] = H;
    }
    return P;
}

// Func_rotate_3Dvol_ROTmat_FourierShear.m:109-130 grids: ifftshifted centred index grids, with every
// voxel on an even-size Nyquist plane (any axis) zeroed in all three grids.
inline double grid_val(int j, int n) { return (double)(((j + n / 2) % n) + 1 - mround((n + 1) / 2.0)); }

// SheerTransforms_for3Drot.m:16-49 on a [z][y][x] complex volume (dims 1,2,3 = x,y,z)
template <typename T> void shear(cx<T>* v, int Nx, int Ny, int Nz, double a, double b, int dim) {
    const int ax1 = dim == 1 ? 2 : dim == 2 ? 2 : 1;   // first fft axis (MATLAB dim), then second
    const int ax2 = dim == 1 ? 1 : dim == 2 ? 0 : 0;
    fft_axis<T>(v, Nx, Ny, Nz, ax1, -1);
    fft_axis<T>(v, Nx, Ny, Nz, ax2, -1);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < Nz; ++z)
        for (int y = 0; y < Ny; ++y)
            for (int x = 0; x < Nx; ++x) {
                double gx = grid_val(x, Nx), gy = grid_val(y, Ny), gz = grid_val(z, Nz);
                const bool nyq = (Nx % 2 == 0 && x == Nx / 2) || (Ny % 2 == 0 && y == Ny / 2) || (Nz % 2 == 0 && z == Nz / 2);
                if (nyq) { gx = gy = gz = 0; }
                double arg;
                if (dim == 1) arg = ((double)Ny / Nx * a * gy / Ny + (double)Nz / Nx * b * gz / Nz) * gx;
                else if (dim == 2) arg = ((double)Nx / Ny * b * gx / Nx + (double)Nz / Ny * a * gz / Nz) * gy;
                else arg = ((double)Nx / Nz * a * gx / Nx + (double)Ny / Nz * b * gy / Ny) * gz;
                const double ph = 2 * M_PI * arg;
                v[((size_t)z * Ny + y) * Nx + x] *= cx<T>((T)std::cos(ph), (T)std::sin(ph));
            }
    fft_axis<T>(v, Nx, Ny, Nz, ax1, +1);
    fft_axis<T>(v, Nx, Ny, Nz, ax2, +1);
}

template <typename T> void rotate(const T* in, T* out, int Nx, int Ny, int Nz, const double MAT[9], double rotStepInit) {
    const ShearPlan P = plan_rotation(MAT, rotStepInit);
    const size_t nv = (size_t)Nx * Ny * Nz;
    std::vector<cx<T>> v(nv);
    for (size_t i = 0; i < nv; ++i) v[i] = cx<T>(in[i], 0);
    shift3<T>(v.data(), Nx, Ny, Nz, true);
    for (int s = 0; s < P.numStep; ++s) {
        for (int k = 0; k < 4; ++k) shear<T>(v.data(), Nx, Ny, Nz, P.ab[k][0], P.ab[k][1], P.dims[k]);
        for (size_t i = 0; i < nv; ++i) v[i] = cx<T>(std::real(v[i]), 0);
    }
    shift3<T>(v.data(), Nx, Ny, Nz, false);
    for (size_t i = 0; i < nv; ++i) out[i] = std::real(v[i]);
}

// ------------------------------------------------------------------------------------------------
// Total-variation regulariser, TV_method.m:5-26 (3D): Chambolle's projection for
// min_u 1/2 |u - f|^2 + lambda TV(u) with periodic forward differences.  Volume [z][y][x], MATLAB dims
// (x, y, z) = (1, 2, 3); circshift(A, [1 0 0]) is A(x - 1), circshift(A, [-1 0 0]) is A(x + 1).
// In MATLAB P starts as double zeros and becomes single after the first update (double op single is
// single), so the float restatement runs in single throughout.
// ------------------------------------------------------------------------------------------------
struct Vol3 {
    int nx, ny, nz;
    size_t at(int x, int y, int z) const {
        x = (x + nx) % nx; y = (y + ny) % ny; z = (z + nz) % nz;
        return ((size_t)z * ny + y) * nx + x;
    }
};

// TV_div.m part: P1 + P2 + P3 - P1(x-1) - P2(y-1) - P3(z-1)
template <typename T> T tv_div(const Vol3& g, const T* P0, const T* P1, const T* P2, int x, int y, int z) {
    const size_t i = g.at(x, y, z);
    T d = P0[i] + P1[i] + P2[i];
    d = d - P0[g.at(x - 1, y, z)];
    d = d - P1[g.at(x, y - 1, z)];
    d = d - P2[g.at(x, y, z - 1)];
    return d;
}

template <typename T>
double tv_method(const Vol3& g, const T* f, T* out, double lambda, int niter) {
    const size_t nv = (size_t)g.nx * g.ny * g.nz;
    std::vector<T> P0(nv, T(0)), P1(nv, T(0)), P2(nv, T(0)), D(nv);
    const T tau = (T)(1.0 / 12.0), lam = (T)lambda;
    for (int it = 0; it < niter; ++it) {
        // temp_term = tau .* TV_del(TV_div(P) - input / lambda)
#pragma omp parallel for schedule(static)
        for (int z = 0; z < g.nz; ++z)
            for (int y = 0; y < g.ny; ++y)
                for (int x = 0; x < g.nx; ++x) {
                    const size_t i = g.at(x, y, z);
                    D[i] = tv_div<T>(g, P0.data(), P1.data(), P2.data(), x, y, z) - f[i] / lam;
                }
        // P = TV_project(P + temp_term)
#pragma omp parallel for schedule(static)
        for (int z = 0; z < g.nz; ++z)
            for (int y = 0; y < g.ny; ++y)
                for (int x = 0; x < g.nx; ++x) {
                    const size_t i = g.at(x, y, z);
                    const T q0 = P0[i] + tau * (D[g.at(x + 1, y, z)] - D[i]);
                    const T q1 = P1[i] + tau * (D[g.at(x, y + 1, z)] - D[i]);
                    const T q2 = P2[i] + tau * (D[g.at(x, y, z + 1)] - D[i]);
                    const T den = std::sqrt(std::max(T(1), q0 * q0 + q1 * q1 + q2 * q2));
                    P0[i] = q0 / den;
                    P1[i] = q1 / den;
                    P2[i] = q2 / den;
                }
    }
    // output = input - lambda * TV_div(P); error = lambda * TV_error(output)
    for (int z = 0; z < g.nz; ++z)
        for (int y = 0; y < g.ny; ++y)
            for (int x = 0; x < g.nx; ++x) {
                const size_t i = g.at(x, y, z);
                out[i] = f[i] - lam * tv_div<T>(g, P0.data(), P1.data(), P2.data(), x, y, z);
            }
    double e = 0.0;
    for (int z = 0; z < g.nz; ++z)
        for (int y = 0; y < g.ny; ++y)
            for (int x = 0; x < g.nx; ++x) {
                const size_t i = g.at(x, y, z);
                const double d0 = (double)(out[g.at(x + 1, y, z)] - out[i]);
                const double d1 = (double)(out[g.at(x, y + 1, z)] - out[i]);
                const double d2 = (double)(out[g.at(x, y, z + 1)] - out[i]);
                e += std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            }
    return lambda * e;
}

// ------------------------------------------------------------------------------------------------
// Position loop (Run_MSET_recon.m:133-158) and error pass (207-222)
// ------------------------------------------------------------------------------------------------
template <typename T>
struct Work {
    std::vector<cx<T>> trans, save, psi0, resi, grad, bw, probe_c, prop_c, back_c;
    Work(const Geo& g) {
        const size_t n2 = (size_t)g.n * g.n;
        trans.resize(n2 * g.L); save.resize(n2 * g.L); psi0.resize(n2); resi.resize(n2);
        grad.resize(n2 * g.nz); bw.resize(n2);
    }
};

// mode bit 2 selects the SSET operator (one plane of depth nz); bit 1 the gradient mode
Geo make_geo(int n, int nz, int bin, double sigma, double dx, int mode = 0) {
    Geo g;
    g.sset = (mode & MODE_SSET) != 0;
    if (g.sset) bin = nz;
    g.n = n; g.nz = nz; g.bin = std::max(1, std::min(bin, nz)); g.L = (nz + g.bin - 1) / g.bin;
    g.sigma = sigma; g.dx = dx;
    return g;
}

template <typename T>
int run_tilt(const Geo& g, int mode, T* V, cx<T>* probe, const cx<T>* prop, const cx<T>* back_prop,
             float* xlist, float* ylist, const float* Mall, int npos, int nsx, int nsy, const double steps[3],
             double* err_acc, double* err_pos) {
    const size_t n2 = (size_t)g.n * g.n;
    Work<T> w(g);
    std::vector<T> M(n2);
    for (int k = 0; k < npos; ++k) {
        const double row = xlist[k], col = ylist[k];
        trans_any<T>(g, V, w.trans.data());
        subpixel_shift<T>(probe, w.psi0.data(), g.n, row, col);
        for (size_t i = 0; i < n2; ++i) M[i] = (T)Mall[(size_t)k * n2 + i];
        const double e = forward_any<T>(g, w.trans.data(), w.psi0.data(), prop, M.data(), w.save.data(), w.resi.data());
        if (err_pos) err_pos[k] = e;
        if (err_acc) *err_acc += e / nsx / nsy;
        double dr, dc;
        back_any<T>(g, mode, w.trans.data(), w.save.data(), w.psi0.data(), w.resi.data(), back_prop, w.grad.data(), w.bw.data(), row, col, &dr, &dc);
        probe_update<T>(probe, w.bw.data(), g.n, row, col, steps[1]);
        xlist[k] = (float)((T)row + (T)steps[2] * (T)dr);
        ylist[k] = (float)((T)col + (T)steps[2] * (T)dc);
        const T s1 = (T)steps[0];
        const size_t nv = n2 * g.nz;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)nv; ++i) V[i] += s1 * std::real(w.grad[i]);
    }
    return 0;
}

template <typename T>
int error_tilt(const Geo& g, const T* V, const cx<T>* probe, const cx<T>* prop, const float* xlist, const float* ylist,
               const float* Mall, int npos, int nsx, int nsy, double* err_acc, double* err_pos) {
    // positions are independent in the error pass (no updates), so they run in parallel; the
    // per-position terms are summed afterwards in position order
    const size_t n2 = (size_t)g.n * g.n;
    std::vector<cx<T>> trans(n2 * g.L);
    trans_any<T>(g, V, trans.data());
    std::vector<double> ep(npos);
#pragma omp parallel
    {
        Work<T> w(g);
        std::vector<T> M(n2);
#pragma omp for schedule(dynamic, 4)
        for (int k = 0; k < npos; ++k) {
            subpixel_shift<T>(probe, w.psi0.data(), g.n, xlist[k], ylist[k]);
            for (size_t i = 0; i < n2; ++i) M[i] = (T)Mall[(size_t)k * n2 + i];
            ep[k] = forward_any<T>(g, trans.data(), w.psi0.data(), prop, M.data(), w.save.data(), nullptr);
        }
    }
    for (int k = 0; k < npos; ++k) {
        if (err_pos) err_pos[k] = ep[k];
        if (err_acc) *err_acc += ep[k] / nsx / nsy;
    }
    return 0;
}

// Gradient sums at a fixed state (the batched path's semantics): over the given positions,
// dV (n2*L, per plane; every voxel of a bin receives the same value) = sum Re(grad),
// dprobe (n2) = sum shift(bw, -d), dpos (2 per position) = position update terms, err per position.
template <typename T>
int grad_fixed(const Geo& g, int mode, const T* V, const cx<T>* probe, const cx<T>* prop, const cx<T>* back_prop,
               const double* xlist, const double* ylist, const float* Mall, int npos,
               T* dV, cx<T>* dprobe, double* dpos, double* err_pos) {
    const size_t n2 = (size_t)g.n * g.n;
    Work<T> w(g);
    std::vector<T> M(n2);
    std::vector<cx<T>> sh(n2);
    trans_any<T>(g, V, w.trans.data());
    std::fill(dV, dV + n2 * g.L, T(0));
    std::fill(dprobe, dprobe + n2, cx<T>(0, 0));
    const double c2 = 2.0 * mround((g.n + 1) / 2.0);
    for (int k = 0; k < npos; ++k) {
        const double row = xlist[k], col = ylist[k];
        subpixel_shift<T>(probe, w.psi0.data(), g.n, row, col);
        for (size_t i = 0; i < n2; ++i) M[i] = (T)Mall[(size_t)k * n2 + i];
        const double e = forward_any<T>(g, w.trans.data(), w.psi0.data(), prop, M.data(), w.save.data(), w.resi.data());
        if (err_pos) err_pos[k] = e;
        double dr, dc;
        back_any<T>(g, mode, w.trans.data(), w.save.data(), w.psi0.data(), w.resi.data(), back_prop, w.grad.data(), w.bw.data(), row, col, &dr, &dc);
        for (int a = 0; a < g.L; ++a)
            for (size_t i = 0; i < n2; ++i) dV[(size_t)a * n2 + i] += std::real(w.grad[(size_t)bin_lo(g, a) * n2 + i]);
        subpixel_shift<T>(w.bw.data(), sh.data(), g.n, c2 - row, c2 - col);
        for (size_t i = 0; i < n2; ++i) dprobe[i] += sh[i];
        dpos[2 * k] = dr;
        dpos[2 * k + 1] = dc;
    }
    return 0;
}

// Amplitude loss at a fixed state, L = sum_k sum_q (|Psi_k| - sqrt(M_k))^2 (for finite differences)
template <typename T>
double loss_fixed(const Geo& g, const T* V, const cx<T>* probe, const cx<T>* prop, const double* xlist,
                  const double* ylist, const float* Mall, int npos) {
    const size_t n2 = (size_t)g.n * g.n;
    Work<T> w(g);
    std::vector<T> M(n2);
    trans_any<T>(g, V, w.trans.data());
    double L = 0;
    for (int k = 0; k < npos; ++k) {
        subpixel_shift<T>(probe, w.psi0.data(), g.n, xlist[k], ylist[k]);
        for (size_t i = 0; i < n2; ++i) M[i] = (T)Mall[(size_t)k * n2 + i];
        L += forward_any<T>(g, w.trans.data(), w.psi0.data(), prop, M.data(), w.save.data(), nullptr) * (double)n2;
    }
    return L;
}

}  // namespace

// ================================================================================================
// C ABI (ctypes).  Complex arrays are interleaved (re, im).  Real-type suffix: _d double, _f float.
// ================================================================================================
extern "C" {

int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

void oracle_fft1_d(double* x, int n, int sign) { fft_axis<double>((cx<double>*)x, n, 1, 1, 0, sign); }
void oracle_fft1_f(float* x, int n, int sign) { fft_axis<float>((cx<float>*)x, n, 1, 1, 0, sign); }
void oracle_fft2_d(double* x, int n, int sign) { fft2<double>((cx<double>*)x, n, sign); }
void oracle_fft2_f(float* x, int n, int sign) { fft2<float>((cx<float>*)x, n, sign); }
void oracle_fft2c_d(double* x, int n, int sign) { fft2c<double>((cx<double>*)x, n, sign); }
void oracle_fft_axis_d(double* x, int n0, int n1, int n2, int axis, int sign) { fft_axis<double>((cx<double>*)x, n0, n1, n2, axis, sign); }

// electron_wavelength.m:1-9 (E0 in eV, lambda in A) and interaction_constant.m:1-10
double oracle_wavelength(double E0) {
    const double m = 9.109383e-31, e = 1.602177e-19, c = 299792458.0, h = 6.62607e-34;
    return h / std::sqrt(2 * m * e * E0) / std::sqrt(1 + e * E0 / 2 / m / c / c) * 1e10;
}
double oracle_sigma(double E0) {
    const double m = 9.109383e-31, e = 1.602177e-19, c = 299792458.0;
    const double lam = oracle_wavelength(E0);
    return (2 * M_PI / lam / E0) * (m * c * c + e * E0) / (2 * m * c * c + e * E0);
}

// Func_generate_propagator.m:17-38, MATLAB (centred) layout [y][x]; back != 0 gives back_prop
void oracle_propagator(double* out, int n, double dx, double lambda, int binning, int back) {
    const double cent = mround((n + 1) / 2.0), mf = 1.0 / (n * dx);
    const double sgn = back ? 1.0 : -1.0;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const double qx = (x + 1 - cent) * mf, qy = (y + 1 - cent) * mf;
            const double ph = sgn * M_PI * lambda * (dx * binning) * (qx * qx + qy * qy);
            const float re = (float)std::cos(ph), im = (float)std::sin(ph);
            out[2 * ((size_t)y * n + x)] = re;
            out[2 * ((size_t)y * n + x) + 1] = im;
        }
}

// Func_generate_probe_wave.m:14-56 for one tilt; ab = {C10, C30, C50, C12, Phi12, C21, Phi21, C23, Phi23}
// (C10 = -probeDefocus, Func_check_probe_wave_parameters.m:136-142).  Output MATLAB layout [y][x].
void oracle_probe(double* out, int n, double dx, double lambda, double alpha_mrad, const double* ab) {
    const double cent = mround((n + 1) / 2.0), mf = 1.0 / (n * dx);
    const double amax = alpha_mrad / 1000.0;
    std::vector<cx<double>> a((size_t)n * n);
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const double qx = (x + 1 - cent) * mf, qy = (y + 1 - cent) * mf;
            const double q2 = qx * qx + qy * qy;
            const double qa = lambda * std::sqrt(q2);
            const double Phi = std::atan2(qy, qx);
            const bool mask = q2 < (amax / lambda) * (amax / lambda);
            double chi = 0.5 * qa * qa * ab[0] + 0.25 * std::pow(qa, 4) * ab[1] + 1.0 / 6 * std::pow(qa, 6) * ab[2]
                         + 0.5 * qa * qa * ab[3] * std::cos(2 * (Phi - ab[4]))
                         + 1.0 / 3 * std::pow(qa, 3) * ab[5] * std::cos(1 * (Phi - ab[6]))
                         + 1.0 / 3 * std::pow(qa, 3) * ab[7] * std::cos(3 * (Phi - ab[8]));
            chi = 2 * M_PI / lambda * chi;
            a[(size_t)y * n + x] = mask ? std::exp(cx<double>(0, -chi)) : cx<double>(0, 0);
        }
    shift2<double>(a.data(), n, n, true);
    fft2<double>(a.data(), n, -1);
    shift2<double>(a.data(), n, n, false);
    std::memcpy(out, a.data(), sizeof(double) * 2 * n * n);
}

// STEP00_INIT.m:214-218: probe *= sqrt(mean_intensity / sum(abs(fft2(probe)).^2))
void oracle_probe_normalise(double* probe, int n, double mean_intensity) {
    std::vector<cx<double>> F((cx<double>*)probe, (cx<double>*)probe + (size_t)n * n);
    fft2<double>(F.data(), n, -1);
    double s = 0;
    for (auto& v : F) s += std::norm(v);
    const double f = std::sqrt(mean_intensity / s);
    for (size_t i = 0; i < (size_t)2 * n * n; ++i) probe[i] *= f;
}

void oracle_subpixel_d(const double* p, double* out, int n, double row, double col) {
    subpixel_shift<double>((const cx<double>*)p, (cx<double>*)out, n, row, col);
}
void oracle_subpixel_f(const float* p, float* out, int n, double row, double col) {
    subpixel_shift<float>((const cx<float>*)p, (cx<float>*)out, n, row, col);
}

void oracle_transmission_d(int n, int nz, int bin, double sigma, double dx, const double* V, double* trans) {
    transmission<double>(make_geo(n, nz, bin, sigma, dx), V, (cx<double>*)trans);
}

// One STEP02_FORWARD_4D from an already-shifted incident wave; returns mean((sqrt M - sqrt I)^2).
double oracle_forward_d(int n, int nz, int bin, double sigma, double dx, const double* V, const double* psi0,
                        const double* prop, const double* M, double* save, double* exitw, double* resi) {
    const Geo g = make_geo(n, nz, bin, sigma, dx);
    std::vector<cx<double>> tr((size_t)n * n * g.L);
    transmission<double>(g, V, tr.data());
    return forward<double>(g, tr.data(), (const cx<double>*)psi0, (const cx<double>*)prop, M, (cx<double>*)save,
                           (cx<double>*)exitw, (cx<double>*)resi);
}

// One STEP03_BACK (object gradient, back wave, position terms) given forward outputs.
void oracle_back_d(int n, int nz, int bin, double sigma, double dx, int mode, const double* V, const double* save,
                   const double* psi0, const double* resi, const double* back_prop, double* grad, double* bw,
                   double row, double col, double* dpos) {
    const Geo g = make_geo(n, nz, bin, sigma, dx);
    std::vector<cx<double>> tr((size_t)n * n * g.L);
    transmission<double>(g, V, tr.data());
    back<double>(g, mode, tr.data(), (const cx<double>*)save, (const cx<double>*)psi0, (const cx<double>*)resi,
                 (const cx<double>*)back_prop, (cx<double>*)grad, (cx<double>*)bw, row, col, &dpos[0], &dpos[1]);
}

int oracle_run_tilt_d(int n, int nz, int bin, double sigma, double dx, int mode, double* V, double* probe,
                      const double* prop, const double* back_prop, float* xlist, float* ylist, const float* M,
                      int npos, int nsx, int nsy, const double* steps, double* err_acc, double* err_pos) {
    return run_tilt<double>(make_geo(n, nz, bin, sigma, dx, mode), mode & 1, V, (cx<double>*)probe, (const cx<double>*)prop,
                            (const cx<double>*)back_prop, xlist, ylist, M, npos, nsx, nsy, steps, err_acc, err_pos);
}
int oracle_run_tilt_f(int n, int nz, int bin, double sigma, double dx, int mode, float* V, float* probe,
                      const float* prop, const float* back_prop, float* xlist, float* ylist, const float* M,
                      int npos, int nsx, int nsy, const double* steps, double* err_acc, double* err_pos) {
    return run_tilt<float>(make_geo(n, nz, bin, sigma, dx, mode), mode & 1, V, (cx<float>*)probe, (const cx<float>*)prop,
                           (const cx<float>*)back_prop, xlist, ylist, M, npos, nsx, nsy, steps, err_acc, err_pos);
}

int oracle_error_tilt_d(int n, int nz, int bin, double sigma, double dx, int mode, const double* V, const double* probe,
                        const double* prop, const float* xlist, const float* ylist, const float* M, int npos,
                        int nsx, int nsy, double* err_acc, double* err_pos) {
    return error_tilt<double>(make_geo(n, nz, bin, sigma, dx, mode), V, (const cx<double>*)probe, (const cx<double>*)prop,
                              xlist, ylist, M, npos, nsx, nsy, err_acc, err_pos);
}
int oracle_error_tilt_f(int n, int nz, int bin, double sigma, double dx, int mode, const float* V, const float* probe,
                        const float* prop, const float* xlist, const float* ylist, const float* M, int npos,
                        int nsx, int nsy, double* err_acc, double* err_pos) {
    return error_tilt<float>(make_geo(n, nz, bin, sigma, dx, mode), V, (const cx<float>*)probe, (const cx<float>*)prop,
                             xlist, ylist, M, npos, nsx, nsy, err_acc, err_pos);
}

int oracle_grad_fixed_d(int n, int nz, int bin, double sigma, double dx, int mode, const double* V, const double* probe,
                        const double* prop, const double* back_prop, const double* xlist, const double* ylist,
                        const float* M, int npos, double* dV, double* dprobe, double* dpos, double* err_pos) {
    return grad_fixed<double>(make_geo(n, nz, bin, sigma, dx, mode), mode & 1, V, (const cx<double>*)probe,
                              (const cx<double>*)prop, (const cx<double>*)back_prop, xlist, ylist, M, npos, dV,
                              (cx<double>*)dprobe, dpos, err_pos);
}

double oracle_loss_d(int n, int nz, int bin, double sigma, double dx, int mode, const double* V, const double* probe,
                     const double* prop, const double* xlist, const double* ylist, const float* M, int npos) {
    return loss_fixed<double>(make_geo(n, nz, bin, sigma, dx, mode), V, (const cx<double>*)probe, (const cx<double>*)prop,
                              xlist, ylist, M, npos);
}

// Func_rot_3Dvol_FourierShear.m:8-12 (inverse = 0) and Func_inv_rot_3Dvol_FourierShear.m:8-13 (inverse = 1):
// R = (R1 R2 R3)' or inv((R1 R2 R3)'); returns 0 if all |angles| <= 0.5 deg (identity), else 1.
int oracle_rot_matrix(const double* v1, const double* v2, const double* v3, double phi, double theta, double psi,
                      int inverse, double* R) {
    double R1[9], R2[9], R3[9], T1[9], T2[9];
    quat_rot(v1, phi, R1);
    quat_rot(v2, theta, R2);
    quat_rot(v3, psi, R3);
    matmul3(R1, R2, T1);
    matmul3(T1, R3, T2);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) R[r * 3 + c] = T2[c * 3 + r];   // transpose
    if (inverse) {
        // rotation matrices are orthogonal: inv(R) = R'
        double Ti[9];
        std::memcpy(Ti, R, sizeof(Ti));
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) R[r * 3 + c] = Ti[c * 3 + r];
    }
    return (std::fabs(theta) > 0.5 || std::fabs(phi) > 0.5 || std::fabs(psi) > 0.5) ? 1 : 0;
}

// Shear plan of Func_rotate_3Dvol_ROTmat_FourierShear for inspection: numStep, dims[4], ab[8]
void oracle_shear_plan(const double* R, double rotStepInit, int* numStep, int* dims, double* ab) {
    const ShearPlan P = plan_rotation(R, rotStepInit);
    *numStep = P.numStep;
    for (int k = 0; k < 4; ++k) { dims[k] = P.dims[k]; ab[2 * k] = P.ab[k][0]; ab[2 * k + 1] = P.ab[k][1]; }
}

// Func_rotate_3Dvol_ROTmat_FourierShear(Vol, R, rotStepInit); volume [z][y][x], MATLAB (Nx, Ny, Nz)
void oracle_rotate_d(const double* in, double* out, int nx, int ny, int nz, const double* R, double rotStepInit) {
    rotate<double>(in, out, nx, ny, nz, R, rotStepInit);
}
void oracle_rotate_f(const float* in, float* out, int nx, int ny, int nz, const double* R, double rotStepInit) {
    rotate<float>(in, out, nx, ny, nz, R, rotStepInit);
}

// TV_method(input, lambda, Niter) for a 3D volume [z][y][x]; returns the TV error term
double oracle_tv_d(const double* in, double* out, int nx, int ny, int nz, double lambda, int niter) {
    return tv_method<double>(Vol3{nx, ny, nz}, in, out, lambda, niter);
}
double oracle_tv_f(const float* in, float* out, int nx, int ny, int nz, double lambda, int niter) {
    return tv_method<float>(Vol3{nx, ny, nz}, in, out, lambda, niter);
}

}  // extern "C"
