// rotate.hpp -- 3D Fourier-shear rotation on the device (Func_rotate_3Dvol_ROTmat_FourierShear.m).
#pragma once
#include <hip/hip_runtime.h>

#include <map>

namespace mset {

struct RotWork {
    float2* w = nullptr;              // complex work volume
    size_t wn = 0;
    std::map<int, float2*> tw;        // twiddle tables by line length
};

struct ShearPlan {
    int num_step = 0;
    int dims[4] = {0, 0, 0, 0};       // shear axis (1 = x, 2 = y, 3 = z) of the four shears of a step
    double ab[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
};

bool rot_size_supported(int n);
// (R1 R2 R3)' (inverse = false) or its inverse; false when every |angle| <= 0.5 (identity)
bool rot_matrix(const double* v1, const double* v2, const double* v3, double phi, double theta, double psi,
                bool inverse, double R[9]);
ShearPlan plan_rotation(const double R[9], double rot_step_init);
// In-place rotation of a device volume float[nz][ny][nx]; returns an MSET_* status.
int rot_run(RotWork& W, float* vol, int nx, int ny, int nz, const double R[9], double rot_step_init, hipStream_t st);
const char* rot_last_error();
void rot_release(RotWork& W);

}  // namespace mset
