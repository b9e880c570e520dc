// advect.hip -- the particle advection sub-step on gfx950.
//
// Reference: src/advection/main_advection.cpp.  The scheme the reference
// compiles (ADVECTION_TYPE 1, :3 and :19-21) is explicit Euler (:44-61):
//     x += u dt,  y += v dt.
// Its "rk2" alternative (:64-83) scales the current velocity only:
//     u1 = u/2, u2 = u1/2, u3 = u2,
//     x += (u + 2 u1 + 2 u2 + u3) dt / 6       (the same for y with v).
// Both are evaluated here in the reference's operation order, and this file
// is compiled without FMA contraction, so each update rounds exactly as the
// reference's host loop does.
//
// One thread per particle; 48 B of HBM traffic per particle (x, y, u, v read,
// x, y written).
#include "ctx.hpp"

namespace vpm {
namespace {

constexpr int kAdvectBlock = 256;

__global__ void __launch_bounds__(kAdvectBlock)
    k_advect(int n, double *__restrict__ x, double *__restrict__ y, const double *__restrict__ u,
             const double *__restrict__ v, double dt, int scheme) {
    const int i = blockIdx.x * kAdvectBlock + threadIdx.x;
    if (i >= n) return;
    const double ui = u[i], vi = v[i];
    if (scheme == kAdvectRk2) {
        const double u1 = ui / 2, u2 = u1 / 2, u3 = u2;
        const double v1 = vi / 2, v2 = v1 / 2, v3 = v2;
        x[i] += (ui + 2 * u1 + 2 * u2 + u3) * dt / 6;
        y[i] += (vi + 2 * v1 + 2 * v2 + v3) * dt / 6;
    } else {
        x[i] += ui * dt;
        y[i] += vi * dt;
    }
}

}  // namespace

void advect(int n, double *x, double *y, const double *u, const double *v, double dt, int scheme,
            hipStream_t s) {
    k_advect<<<div_up(n, kAdvectBlock), kAdvectBlock, 0, s>>>(n, x, y, u, v, dt, scheme);
    VPM_LAUNCH_CHECK();
}

}  // namespace vpm
