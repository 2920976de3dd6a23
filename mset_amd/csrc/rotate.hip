// rotate.hip -- 3D rotation by Fourier shears on the device.
//
// Restates Func_rotate_3Dvol_ROTmat_FourierShear.m:14-133 (Welling et al., Graphical Models 68 (2006)
// 356-370) with its helpers My_obtain_quaternion_from_matrix.m, calculate_S_matrices_for3DFourierShear
// Rot_perm{,_backward}_ver1.m and SheerTransforms_for3Drot.m, plus the matrices of
// Func_rot_3Dvol_FourierShear.m / Func_inv_rot_3Dvol_FourierShear.m and MatrixQuaternionRot.m.
//
// A rotation is numStep repetitions of four shears.  Each shear is a 2D transform over the two axes
// other than its shear axis, a phase ramp, and the inverse transform; the real part is taken after
// the fourth shear of every step.  Transforms along different axes commute and nothing but the phase
// ramps happens between shears, so the engine tracks which axes are currently in the Fourier domain
// and only transforms the axes whose domain must change: a step costs 10 one-axis transforms instead
// of 16, and each shear's phase ramp is applied right after the transform that completes its domain
// change.  The transforms of one domain change may run in any order; they are ordered so that a shear
// starts on the axis the previous one ended on, and those two transforms share one pass over the
// volume (7 passes per step instead of 10 for the shear axes of a tilt about y).
//
// A pass is a batched 1D FFT along one axis of the complex work volume, or two FFTs of the same lines:
// each workgroup loads a tile of TL lines into LDS (for the y and z axes a tile is TL consecutive x
// positions, so loads stay coalesced), runs the Stockham stages of fft_lds.hpp's radix plans with
// consecutive lanes on consecutive lines (so LDS accesses are free of bank conflicts), and writes the
// tile back with the 1/N of an inverse transform, the phase ramp and the real part fused into the store
// (between two transforms, into an LDS sweep).
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/mset.h"
#include "fft_lds.hpp"
#include "rotate.hpp"

namespace mset {

namespace {
thread_local std::string r_err;
int rfail(int code, const std::string& m) {
    r_err = m;
    return code;
}
}  // namespace

const char* rot_last_error() { return r_err.c_str(); }

// ------------------------------------------------------------------ device side
// One transform of the lines and the element-wise work that follows it
struct LineOp {
    int inv;            // 1: inverse transform (with 1/N)
    int phase_dim;      // 0: none; 1, 2, 3: shear along x, y, z with coefficients (a, b)
    double a, b;
    int take_real;      // keep the real part only
};

struct PassArgs {
    float2* w;
    int nx, ny, nz;
    int nops;           // 1, or 2: a second transform of the same lines while the tile is in LDS
    LineOp op[2];
    const float2* tw;
};

// index grid of the shear phase (Func_rotate_3Dvol_ROTmat_FourierShear.m:109-113): ifftshifted
// (1:n) - round((n+1)/2) at unshifted index j
__device__ __forceinline__ double shear_grid(int j, int n) {
    return (double)(((j + n / 2) % n) + 1 - (int)floor((n + 1) / 2.0 + 0.5));
}

__device__ __forceinline__ float2 shear_phase(const PassArgs& P, const LineOp& o, int jx, int jy, int jz) {
    // even-size Nyquist planes of any axis zero all three grids (:115-129) -> phase 1
    if ((P.nx % 2 == 0 && jx == P.nx / 2) || (P.ny % 2 == 0 && jy == P.ny / 2) || (P.nz % 2 == 0 && jz == P.nz / 2))
        return make_float2(1.f, 0.f);
    const double gx = shear_grid(jx, P.nx), gy = shear_grid(jy, P.ny), gz = shear_grid(jz, P.nz);
    double arg;   // SheerTransforms_for3Drot.m:29-48, exp(1i*2*pi*arg)
    if (o.phase_dim == 1) arg = (o.a * gy + o.b * gz) * gx / P.nx;
    else if (o.phase_dim == 2) arg = (o.b * gx + o.a * gz) * gy / P.ny;
    else arg = (o.a * gx + o.b * gy) * gz / P.nz;
    const float f = (float)(arg - rint(arg));
    float sn, cs;
    sincospif(2.f * f, &sn, &cs);
    return make_float2(cs, sn);
}

template <int N, int AXIS, int TL, int NT>
__global__ __launch_bounds__(NT) void k_axis_pass(PassArgs P) {
    // x lines are stored one per LDS row with an odd pitch (N + 1); y / z tiles as [element][line]
    constexpr int ES = AXIS == 0 ? 1 : TL;
    constexpr int LS = AXIS == 0 ? N + 1 : 1;
    __shared__ float2 s[AXIS == 0 ? (N + 1) * TL : N * TL];
    __shared__ float2 tw[N];
    for (int i = threadIdx.x; i < N; i += NT) tw[i] = P.tw[i];
    const size_t sxy = (size_t)P.nx * P.ny;
    // Tile origin.  Workgroups go to the 8 XCDs round-robin (XCD = blockIdx % 8); for the x and y passes
    // the tiles are renumbered so that XCD c always works on the same eighth of the z planes, whose data
    // then stays in that XCD's L2 from one pass to the next (measured: 15 % less rotation time at
    // 128 x 128 x 64).  A z pass reads across all planes and keeps the tile order, which measured
    // faster than splitting it by y.
    constexpr int XCD = 8;
    int b = blockIdx.x;
    if (AXIS != 2 && gridDim.x % XCD == 0) b = (b % XCD) * (gridDim.x / XCD) + b / XCD;
    int row0 = 0, x0 = 0, fixed = 0;
    if (AXIS == 0) row0 = b * TL;
    else {
        const int xt = (P.nx + TL - 1) / TL;
        fixed = b / xt;                     // z for AXIS 1, y for AXIS 2
        x0 = (b % xt) * TL;
    }
    const int nrows = P.ny * P.nz;
    auto gidx = [&](int l, int e, bool& ok, int& jx, int& jy, int& jz) -> size_t {
        if (AXIS == 0) {
            const int row = row0 + l;
            ok = row < nrows;
            jx = e; jy = row % P.ny; jz = row / P.ny;
            return (size_t)row * P.nx + e;
        } else if (AXIS == 1) {
            jx = x0 + l; jy = e; jz = fixed;
            ok = jx < P.nx;
            return (size_t)jz * sxy + (size_t)jy * P.nx + jx;
        } else {
            jx = x0 + l; jy = fixed; jz = e;
            ok = jx < P.nx;
            return (size_t)jz * sxy + (size_t)jy * P.nx + jx;
        }
    };
    for (int i = threadIdx.x; i < N * TL; i += NT) {
        const int l = AXIS == 0 ? i / N : i % TL, e = AXIS == 0 ? i % N : i / TL;
        bool ok;
        int jx, jy, jz;
        const size_t g = gidx(l, e, ok, jx, jy, jz);
        s[l * LS + e * ES] = ok ? P.w[g] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    // the inverse transform's 1/N, the shear phase and the real part after a transform
    auto finish = [&](const LineOp& o, float2 v, int jx, int jy, int jz) {
        if (o.inv) v = cscale(v, 1.0f / (float)N);
        if (o.phase_dim) v = cmul(v, shear_phase(P, o, jx, jy, jz));
        if (o.take_real) v.y = 0.f;
        return v;
    };
    if (P.op[0].inv) line_fft<N, NT, TL, ES, LS, 0, 1, true>(s, tw);
    else line_fft<N, NT, TL, ES, LS, 0, 1, false>(s, tw);
    if (P.nops > 1) {
        for (int i = threadIdx.x; i < N * TL; i += NT) {
            const int l = AXIS == 0 ? i / N : i % TL, e = AXIS == 0 ? i % N : i / TL;
            bool ok;
            int jx, jy, jz;
            (void)gidx(l, e, ok, jx, jy, jz);
            if (ok) s[l * LS + e * ES] = finish(P.op[0], s[l * LS + e * ES], jx, jy, jz);
        }
        __syncthreads();
        if (P.op[1].inv) line_fft<N, NT, TL, ES, LS, 0, 1, true>(s, tw);
        else line_fft<N, NT, TL, ES, LS, 0, 1, false>(s, tw);
    }
    const LineOp last = P.nops > 1 ? P.op[1] : P.op[0];
    for (int i = threadIdx.x; i < N * TL; i += NT) {
        const int l = AXIS == 0 ? i / N : i % TL, e = AXIS == 0 ? i % N : i / TL;
        bool ok;
        int jx, jy, jz;
        const size_t g = gidx(l, e, ok, jx, jy, jz);
        if (!ok) continue;
        P.w[g] = finish(last, s[l * LS + e * ES], jx, jy, jz);
    }
}

// w = ifftshift(vol) as complex (Func_rotate_3Dvol_ROTmat_FourierShear.m:132-133)
__global__ void k_rot_load(const float* __restrict__ vol, float2* __restrict__ w, int nx, int ny, int nz) {
    const size_t tot = (size_t)nx * ny * nz;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((size_t)nx * ny));
        const size_t src = ((size_t)((z + nz / 2) % nz) * ny + (y + ny / 2) % ny) * nx + (x + nx / 2) % nx;
        w[i] = make_float2(vol[src], 0.f);
    }
}

// vol = fftshift(real(w)) (:169)
__global__ void k_rot_store(const float2* __restrict__ w, float* __restrict__ vol, int nx, int ny, int nz) {
    const size_t tot = (size_t)nx * ny * nz;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((size_t)nx * ny));
        const size_t src = ((size_t)((z + nz - nz / 2) % nz) * ny + (y + ny - ny / 2) % ny) * nx + (x + nx - nx / 2) % nx;
        vol[i] = w[src].x;
    }
}

// ------------------------------------------------------------------ host side
bool rot_size_supported(int n) {
    return n == 16 || n == 24 || n == 32 || n == 48 || n == 64 || n == 96 || n == 128 || n == 256;
}

namespace {

template <int N, int AXIS> hipError_t launch_pass_n(const PassArgs& P, hipStream_t st) {
    constexpr int TL = N >= 256 ? 8 : 16;
    constexpr int NT = 256;
    int grid;
    if (AXIS == 0) grid = (P.ny * P.nz + TL - 1) / TL;
    else grid = ((P.nx + TL - 1) / TL) * (AXIS == 1 ? P.nz : P.ny);
    hipLaunchKernelGGL((k_axis_pass<N, AXIS, TL, NT>), dim3(grid), dim3(NT), 0, st, P);
    return hipGetLastError();
}

template <int AXIS> hipError_t launch_pass(int n, const PassArgs& P, hipStream_t st) {
    switch (n) {
        case 16: return launch_pass_n<16, AXIS>(P, st);
        case 24: return launch_pass_n<24, AXIS>(P, st);
        case 32: return launch_pass_n<32, AXIS>(P, st);
        case 48: return launch_pass_n<48, AXIS>(P, st);
        case 64: return launch_pass_n<64, AXIS>(P, st);
        case 96: return launch_pass_n<96, AXIS>(P, st);
        case 128: return launch_pass_n<128, AXIS>(P, st);
        case 256: return launch_pass_n<256, AXIS>(P, st);
        default: return hipErrorInvalidValue;
    }
}

// MatrixQuaternionRot.m:1-11 (row-major 3x3)
void axis_angle(const double* v, double deg, double M[9]) {
    const double th = deg * M_PI / 180.0;
    const double nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double h = std::sin(th / 2);
    const double w = std::cos(th / 2), x = -h * v[0] / nv, y = -h * v[1] / nv, z = -h * v[2] / nv;
    M[0] = 1 - 2 * y * y - 2 * z * z; M[1] = 2 * x * y + 2 * w * z;     M[2] = 2 * x * z - 2 * w * y;
    M[3] = 2 * x * y - 2 * w * z;     M[4] = 1 - 2 * x * x - 2 * z * z; M[5] = 2 * y * z + 2 * w * x;
    M[6] = 2 * x * z + 2 * w * y;     M[7] = 2 * y * z - 2 * w * x;     M[8] = 1 - 2 * x * x - 2 * y * y;
}

// My_obtain_quaternion_from_matrix.m:11-39 (returns x, y, z, w)
void matrix_quaternion(const double M[9], double q[4]) {
    const double tr = M[0] + M[4] + M[8];
    if (tr > 0) {
        const double S = 2 * std::sqrt(tr + 1);
        q[3] = S / 4; q[0] = (M[7] - M[5]) / S; q[1] = (M[2] - M[6]) / S; q[2] = (M[3] - M[1]) / S;
    } else if (M[0] > M[4] && M[0] > M[8]) {
        const double S = 2 * std::sqrt(1 + M[0] - M[4] - M[8]);
        q[3] = (M[7] - M[5]) / S; q[0] = S / 4; q[1] = (M[1] + M[3]) / S; q[2] = (M[2] + M[6]) / S;
    } else if (M[4] > M[8]) {
        const double S = 2 * std::sqrt(1 + M[4] - M[0] - M[8]);
        q[3] = (M[2] - M[6]) / S; q[0] = (M[1] + M[3]) / S; q[1] = S / 4; q[2] = (M[5] + M[7]) / S;
    } else {
        const double S = 2 * std::sqrt(1 + M[8] - M[0] - M[4]);
        q[3] = (M[3] - M[1]) / S; q[0] = (M[2] + M[6]) / S; q[1] = (M[5] + M[7]) / S; q[2] = S / 4;
    }
}

// calculate_S_matrices_for3DFourierShearRot_perm{,_backward}_ver1.m: coefficients a..h of the
// decomposition for the cyclic permutation `perm` of (x, y, z); backward negates w and the result.
// Returns the sum of |coefficients|, or +inf where the decomposition is singular.
double decompose(const double q[4], int perm, bool backward, double c[8]) {
    const double X = q[perm % 3], Y = q[(perm + 1) % 3], Z = q[(perm + 2) % 3];
    const double W = backward ? -q[3] : q[3];
    const double d1 = Y * Z - X * W, d2 = X * Y - Z * W;
    double v[8];
    if (std::fabs(d1) > 1e-5 && std::fabs(d2) > 1e-5) {
        v[0] = (X * X + Y * Y) / d1;
        v[1] = (Y * Y - Z * Z) / d2 - 2 * (Y * Z * (X * X + Y * Y)) / d2 / d1;
        v[2] = 2 * Y * Z / d2;
        v[3] = 2 * (X * W - Y * Z);
        v[4] = 2 * (X * Y - Z * W);
        v[5] = -2 * X * Y / d1;
        v[6] = (X * X - Y * Y) / d1 + 2 * X * Y * (Y * Y + Z * Z) / d2 / d1;
        v[7] = (-1 + X * X + W * W) / d2;
    } else if (std::fabs(Y) < 1e-6) {
        v[0] = -X / W; v[1] = Z / W; v[2] = 0; v[3] = 2 * X * W; v[4] = -2 * Z * W; v[5] = 0; v[6] = -X / W; v[7] = Z / W;
    } else {
        return INFINITY;
    }
    double s = 0;
    for (int i = 0; i < 8; ++i) {
        s += std::fabs(v[i]);
        c[i] = backward ? -v[i] : v[i];
    }
    return s;
}

}  // namespace

bool rot_matrix(const double* v1, const double* v2, const double* v3, double phi, double theta, double psi,
                bool inverse, double R[9]) {
    double A[9], B[9], C[9], AB[9], ABC[9];
    axis_angle(v1, phi, A);
    axis_angle(v2, theta, B);
    axis_angle(v3, psi, C);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            AB[r * 3 + c] = A[r * 3] * B[c] + A[r * 3 + 1] * B[3 + c] + A[r * 3 + 2] * B[6 + c];
        }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            ABC[r * 3 + c] = AB[r * 3] * C[c] + AB[r * 3 + 1] * C[3 + c] + AB[r * 3 + 2] * C[6 + c];
        }
    // (R1 R2 R3)'  and, for the inverse rotation, inv((R1 R2 R3)') = R1 R2 R3 (orthogonal)
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) R[r * 3 + c] = inverse ? ABC[r * 3 + c] : ABC[c * 3 + r];
    return std::fabs(theta) > 0.5 || std::fabs(phi) > 0.5 || std::fabs(psi) > 0.5;
}

ShearPlan plan_rotation(const double R[9], double rot_step_init) {
    double q[4];
    matrix_quaternion(R, q);
    const double ang = 2.0 * std::acos(q[3]) * 180.0 / M_PI;
    const double nq = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    ShearPlan P;
    P.num_step = (int)std::floor(ang / rot_step_init) + 1;
    const double half = ang / P.num_step / 2.0 * M_PI / 180.0;
    const double qs[4] = {-std::sin(half) * q[0] / nq, -std::sin(half) * q[1] / nq, -std::sin(half) * q[2] / nq,
                          std::cos(half)};
    double cf[3], cb[3], tmp[8];
    for (int i = 1; i <= 3; ++i) {
        cf[i - 1] = decompose(qs, i, false, tmp);
        cb[i - 1] = decompose(qs, i, true, tmp);
    }
    int f = 0, b = 0;
    for (int i = 1; i < 3; ++i) {
        if (cf[i] < cf[f]) f = i;
        if (cb[i] < cb[b]) b = i;
    }
    const bool fwd = cf[f] < cb[b];
    const int m = (fwd ? f : b) + 1;
    double c[8];
    decompose(qs, m, !fwd, c);
    // shear order and dimensions (Func_rotate_3Dvol_ROTmat_FourierShear.m:155-166)
    const int d_first = (m + 1) % 3 + 1, d_mid1 = fwd ? m % 3 + 1 : (m + 2) % 3 + 1, d_mid2 = fwd ? (m + 2) % 3 + 1 : m % 3 + 1;
    const int order_fwd[4] = {6, 4, 2, 0}, order_bwd[4] = {0, 2, 4, 6};
    const int dims[4] = {d_first, d_mid1, d_mid2, d_first};
    for (int k = 0; k < 4; ++k) {
        const int o = fwd ? order_fwd[k] : order_bwd[k];
        P.dims[k] = dims[k];
        P.ab[k][0] = c[o];
        P.ab[k][1] = c[o + 1];
    }
    return P;
}

static float2* twiddle_table(RotWork& W, int n, hipStream_t st, hipError_t& e) {
    auto it = W.tw.find(n);
    if (it != W.tw.end()) return it->second;
    std::vector<float2> h(n);
    for (int k = 0; k < n; ++k) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        h[k] = make_float2((float)cosl(a), (float)sinl(a));
    }
    float2* d = nullptr;
    e = hipMalloc((void**)&d, sizeof(float2) * n);
    if (e != hipSuccess) return nullptr;
    e = hipMemcpyAsync(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    W.tw[n] = d;
    return d;
}

namespace {

struct AxisOp {
    int axis;
    LineOp op;
};

// The axes one domain change transforms (in any order) and what follows the change: the shear phase of a
// shear, or the real part at the end of a step.  Both belong to the last transform of the group.
struct LineGroup {
    int n = 0;
    int axis[3];
    bool inv[3];
    int phase_dim = 0;
    double a = 0, b = 0;
    bool take_real = false;
};

// Func_rotate_3Dvol_ROTmat_FourierShear.m:155-166: every step runs its four shears, each a transform over
// the two axes other than its shear axis, then returns to real space and keeps the real part.  Transforms
// along different axes commute, so a shear only transforms the axes whose domain must change.
bool shear_groups(const ShearPlan& P, std::vector<LineGroup>& out) {
    bool fourier[3] = {false, false, false};
    for (int s = 0; s < P.num_step; ++s) {
        for (int k = 0; k < 4; ++k) {
            LineGroup g;
            const int S = P.dims[k] - 1;
            for (int ax = 0; ax < 3; ++ax)
                if (fourier[ax] != (ax != S)) {
                    g.axis[g.n] = ax;
                    g.inv[g.n++] = fourier[ax];
                    fourier[ax] = !fourier[ax];
                }
            if (g.n == 0) return false;
            g.phase_dim = P.dims[k];
            g.a = P.ab[k][0];
            g.b = P.ab[k][1];
            out.push_back(g);
        }
        LineGroup g;   // back to real space, real part (:160, :165)
        for (int ax = 0; ax < 3; ++ax)
            if (fourier[ax]) {
                g.axis[g.n] = ax;
                g.inv[g.n++] = true;
                fourier[ax] = false;
            }
        g.take_real = true;
        if (g.n) out.push_back(g);
    }
    return true;
}

// Orders the transforms of each group so that as many groups as possible start on the axis the previous
// group ended on (a dynamic programme over the last axis); such neighbours share one pass over the volume.
// With the shear axes (z, x, y, z) of a tilt about y, a step takes 7 passes instead of 10.
std::vector<AxisOp> order_ops(const std::vector<LineGroup>& G) {
    const int ng = (int)G.size();
    if (ng == 0) return {};
    static const int perm2[2][2] = {{0, 1}, {1, 0}};
    static const int perm3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    auto nperm = [&](int g) { return G[g].n == 1 ? 1 : G[g].n == 2 ? 2 : 6; };
    auto at = [&](int g, int p, int i) { return G[g].n == 1 ? 0 : G[g].n == 2 ? perm2[p][i] : perm3[p][i]; };
    // best[g][a]: most shared passes over groups 0..g when group g ends on axis a; order[g][a]: the order of
    // group g that does it; prev[g][a]: the axis group g - 1 then ends on
    std::vector<std::array<int, 3>> best(ng), order(ng), prev(ng);
    for (int g = 0; g < ng; ++g) {
        best[g] = {-1, -1, -1};
        for (int p = 0; p < nperm(g); ++p) {
            const int first = G[g].axis[at(g, p, 0)], last = G[g].axis[at(g, p, G[g].n - 1)];
            int score = 0, from = -1;
            if (g > 0) {
                score = -1;
                for (int a = 0; a < 3; ++a)
                    if (best[g - 1][a] >= 0 && best[g - 1][a] + (a == first) > score) {
                        score = best[g - 1][a] + (a == first);
                        from = a;
                    }
            }
            if (score > best[g][last]) {
                best[g][last] = score;
                order[g][last] = p;
                prev[g][last] = from;
            }
        }
    }
    std::vector<int> perm(ng);
    int a = 0;
    for (int b = 1; b < 3; ++b)
        if (best[ng - 1][b] > best[ng - 1][a]) a = b;
    for (int g = ng - 1; g >= 0; --g) {
        perm[g] = order[g][a];
        a = prev[g][a];
    }
    std::vector<AxisOp> ops;
    for (int g = 0; g < ng; ++g)
        for (int i = 0; i < G[g].n; ++i) {
            const int j = at(g, perm[g], i);
            const bool last = i == G[g].n - 1;
            AxisOp o;
            o.axis = G[g].axis[j];
            o.op.inv = G[g].inv[j] ? 1 : 0;
            o.op.phase_dim = last ? G[g].phase_dim : 0;
            o.op.a = last ? G[g].a : 0.0;
            o.op.b = last ? G[g].b : 0.0;
            o.op.take_real = last && G[g].take_real ? 1 : 0;
            ops.push_back(o);
        }
    return ops;
}

}  // namespace

int rot_run(RotWork& W, float* vol, int nx, int ny, int nz, const double R[9], double rot_step_init, hipStream_t st) {
    if (!rot_size_supported(nx) || !rot_size_supported(ny) || !rot_size_supported(nz))
        return rfail(MSET_EUNSUPPORTED, "rotation line length not built");
    const size_t nv = (size_t)nx * ny * nz;
    if (W.wn < nv) {
        if (W.w) (void)hipFree(W.w);
        W.w = nullptr;
        W.wn = 0;
        hipError_t e = hipMalloc((void**)&W.w, sizeof(float2) * nv);
        if (e != hipSuccess) return rfail(MSET_ENOMEM, std::string("rotation work volume: ") + hipGetErrorString(e));
        W.wn = nv;
    }
    const int len[3] = {nx, ny, nz};
    const float2* tw[3];
    for (int a = 0; a < 3; ++a) {
        hipError_t e = hipSuccess;
        tw[a] = twiddle_table(W, len[a], st, e);
        if (!tw[a]) return rfail(MSET_EDEVICE, std::string("twiddles: ") + hipGetErrorString(e));
    }
    const ShearPlan P = plan_rotation(R, rot_step_init);
    const unsigned g = (unsigned)std::min<size_t>((nv + 255) / 256, 8192);
    hipLaunchKernelGGL(k_rot_load, dim3(g), dim3(256), 0, st, vol, W.w, nx, ny, nz);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return rfail(MSET_EDEVICE, std::string("rotation load: ") + hipGetErrorString(e));

    std::vector<LineGroup> groups;
    if (!shear_groups(P, groups)) return rfail(MSET_EDEVICE, "shear plan repeats an axis");
    const std::vector<AxisOp> ops = order_ops(groups);
    for (size_t i = 0; i < ops.size();) {
        // a transform of the same lines right after this one runs in the same pass
        const bool two = i + 1 < ops.size() && ops[i + 1].axis == ops[i].axis;
        PassArgs A{};
        A.w = W.w; A.nx = nx; A.ny = ny; A.nz = nz;
        A.nops = two ? 2 : 1;
        A.op[0] = ops[i].op;
        if (two) A.op[1] = ops[i + 1].op;
        const int axis = ops[i].axis;
        A.tw = tw[axis];
        e = axis == 0 ? launch_pass<0>(nx, A, st) : axis == 1 ? launch_pass<1>(ny, A, st) : launch_pass<2>(nz, A, st);
        if (e != hipSuccess) return rfail(MSET_EDEVICE, std::string("shear pass: ") + hipGetErrorString(e));
        i += two ? 2 : 1;
    }
    hipLaunchKernelGGL(k_rot_store, dim3(g), dim3(256), 0, st, W.w, vol, nx, ny, nz);
    e = hipGetLastError();
    if (e != hipSuccess) return rfail(MSET_EDEVICE, std::string("rotation store: ") + hipGetErrorString(e));
    return MSET_OK;
}

void rot_release(RotWork& W) {
    if (W.w) (void)hipFree(W.w);
    W.w = nullptr;
    W.wn = 0;
    for (auto& kv : W.tw) (void)hipFree(kv.second);
    W.tw.clear();
}

}  // namespace mset
