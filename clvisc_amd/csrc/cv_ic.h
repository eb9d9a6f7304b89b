// cv_ic.h -- initial conditions on the device: the optical Glauber model of
// pyvisc/kernel/kernel_glauber.cl (glauber_ini :82-102, num_of_binary_collisions :104-121).
// FP32 like the reference; the constants and the quadrature are the model's definition.
#pragma once
#include "cv_device.h"
#include "../../include/clvisc_amd.h"

namespace cv {

// Nuclear thickness T(x, y) = 2 int_0^{5R} rho_WS(sqrt(x^2 + y^2 + z^2)) dz by five-point
// Gauss-Legendre on ten equal sub-intervals (kernel_glauber.cl:10-33).
__device__ __forceinline__ float gl_thickness(const cv_glauber &g, float x, float y) {
    const float zk[5] = {-0.9061798f, -0.5386492f, 0.0f, 0.5386493f, 0.9061798f};
    const float Ak[5] = {0.2369269f, 0.4786287f, 0.5688889f, 0.4786287f, 0.2369269f};
    const float cut = 5.0f * g.ra;
    float t = 0.0f;
    for (int j = 0; j < 10; j++) {
        const float a = (float)j * cut / 10.0f, b = (float)(j + 1) * cut / 10.0f;
        for (int i = 0; i < 5; i++) {
            const float z = (a + b) / 2.0f + (b - a) / 2.0f * zk[i];
            const float r = sqrtf(x * x + y * y + z * z);
            const float f = g.ro0 / (1.0f + expf((r - g.ra) / g.eta));
            t = t + Ak[i] * f * (b - a) / 2.0f;
        }
    }
    return 2.0f * t;
}

// wounded nucleons and binary collisions (kernel_glauber.cl:72-81)
__device__ __forceinline__ float gl_nw(const cv_glauber &g, float ta, float tb) {
    return ta * (1.0f - powf(1.0f - g.si0 * tb / g.num_nucleons, g.num_nucleons)) +
           tb * (1.0f - powf(1.0f - g.si0 * ta / g.num_nucleons, g.num_nucleons));
}
__device__ __forceinline__ float gl_nb(const cv_glauber &g, float ta, float tb) { return g.si0 * ta * tb; }
__device__ __forceinline__ float gl_ed_transverse(const cv_glauber &g, float ta, float tb) {
    return g.hwn * gl_nw(g, ta, tb) + (1.0f - g.hwn) * gl_nb(g, ta, tb);
}
// plateau with Gaussian tails along eta_s (kernel_glauber.cl:36-48; etas0 = 0 as there)
__device__ __forceinline__ float gl_weight_eta(const cv_glauber &g, float z) {
    if (g.bjorken) return 1.0f;
    const float d = fabsf(z - 0.0f);
    if (d > g.eta_flat) return expf(-powf(d - g.eta_flat, 2.0f) / (2.0f * g.eta_gw * g.eta_gw));
    return 1.0f;
}

// one thread per (i, j): e(x, y) H(eta) into the four planes of ev (velocities zero)
__global__ __launch_bounds__(256) void k_glauber_ini(PhysP p, cv_glauber g, float *ev) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)p.nx * p.ny) return;
    const int i = (int)(t / p.ny), j = (int)(t - (int64_t)(t / p.ny) * p.ny);
    const float x = (float)(i - p.nx / 2) * p.dx;
    const float y = (float)(j - p.ny / 2) * p.dy;
    const float tcent = gl_thickness(g, 0.0f, 0.0f);
    const float ed_central = gl_ed_transverse(g, tcent, tcent);
    const float kfac = g.edmax / ed_central;
    const float b = g.b;
    const float ta = gl_thickness(g, x - 0.5f * b, y);
    const float tb = gl_thickness(g, x + 0.5f * b, y);
    const float edxy = kfac * gl_ed_transverse(g, ta, tb);
    const int64_t n = p.n;
    for (int k = 0; k < p.nz; k++) {
        const float etas = (float)(k - p.nz / 2) * p.dz;
        const int64_t I = t * p.nz + k;
        ev[I] = edxy * gl_weight_eta(g, etas);
        ev[n + I] = 0.0f;
        ev[2 * n + I] = 0.0f;
        ev[3 * n + I] = 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_glauber_nbinary(cv_glauber g, float xlow, float ylow, float dx, float dy,
                                                         int nx, int ny, float *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int i = t / ny, j = t - (t / ny) * ny;
    const float x = xlow + (float)i * dx, y = ylow + (float)j * dy;
    const float ta = gl_thickness(g, x - 0.5f * g.b, y);
    const float tb = gl_thickness(g, x + 0.5f * g.b, y);
    out[t] = gl_nb(g, ta, tb);
}

}  // namespace cv
