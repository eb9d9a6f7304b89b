// cv_vort.h -- vorticity on the grid (pyvisc/kernel/kernel_vorticity.cl: omega :38-99,
// omega_shear :143-221, omega_accT :225-337, omega_mu :103-140), all of a cell's tensors in
// one pass over its stencil.  Output: 32 SoA planes per slot,
//   0..5   thermal vorticity omega^{tx, ty, tz, xy, xz, yz} (hbar c / 2 eps d(beta u))
//   6..21  omega_shear1: (1/2) d_mu u_nu for mu in (tau, x, y, eta), nu = 0..3 (16)
//   22..25 omega_shear2: (1/2) u^mu d_mu u_nu (4)
//   26..31 omega_accT: the acceleration-temperature-gradient vorticity (6)
// Same arithmetic contract as the hydro kernels (a/b as a RN(1/b), no contraction).  An
// axis with a single cell has zero partial derivative (the reference would read past the
// grid); the eta Christoffel terms stay.
#pragma once
#include "cv_device.h"

namespace cv {

#define CV_VORT_N 32

struct Vq {
    float4 ub;      // beta u_mu = g_{mu mu} u^mu / T
    float4 uk;      // u_mu
    float T;
};

template <int EK>
__device__ __forceinline__ Vq vort_q(const EosT<EK> &eos, const float *__restrict__ ev, int64_t n, int64_t I) {
    const float e = ev[I], vx = ev[n + I], vy = ev[2 * n + I], vz = ev[3 * n + I];
    const float g = gamma3(vx, vy, vz);
    const float u0 = g * 1.0f, u1 = g * vx, u2 = g * vy, u3 = g * vz;
    const float T = eos.T(e);
    const float rT = rcp(T);
    Vq q;
    q.uk = make_float4(1.0f * u0, -1.0f * u1, -1.0f * u2, -1.0f * u3);
    q.ub = make_float4(q.uk.x * rT, q.uk.y * rT, q.uk.z * rT, q.uk.w * rT);
    q.T = T;
    return q;
}

__device__ __forceinline__ float4 sub4(float4 a, float4 b) { return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
__device__ __forceinline__ float4 mul4(float4 a, float r) { return make_float4(a.x * r, a.y * r, a.z * r, a.w * r); }

template <int EK>
__global__ __launch_bounds__(256) void k_vorticity(PhysP p, EosP E, const float *__restrict__ ev_old,
                                                   const float *__restrict__ ev_new, float tau, float *om) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= p.n) return;
    EosT<EK> eos(E, E.tab);
    const int64_t n = p.n;
    const int k = (int)(I % p.nz);
    const int j = (int)((I / p.nz) % p.ny);
    const int i = (int)(I / ((int64_t)p.nz * p.ny));
    const Vq qo = vort_q(eos, ev_old, n, I), qn = vort_q(eos, ev_new, n, I);
    // derivative along one axis of the new field: central inside, one-sided at the edges
    float4 dub[3], duk[3];
    float dT[3];
    const int idx[3] = {i, j, k}, dim[3] = {p.nx, p.ny, p.nz};
    const int64_t stride[3] = {(int64_t)p.ny * p.nz, (int64_t)p.nz, 1};
    const float r2h[3] = {rcp(2.0f * p.dx), rcp(2.0f * p.dy), rcp(2.0f * p.dz * tau)};
    const float r1h[3] = {rcp(p.dx), rcp(p.dy), rcp(p.dz * tau)};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (dim[a] == 1) {
            dub[a] = make_float4(0.f, 0.f, 0.f, 0.f); duk[a] = dub[a]; dT[a] = 0.0f;
        } else if (idx[a] != 0 && idx[a] != dim[a] - 1) {
            const Vq qp = vort_q(eos, ev_new, n, I + stride[a]), qm = vort_q(eos, ev_new, n, I - stride[a]);
            dub[a] = mul4(sub4(qp.ub, qm.ub), r2h[a]);
            duk[a] = mul4(sub4(qp.uk, qm.uk), r2h[a]);
            dT[a] = (qp.T - qm.T) * r2h[a];
        } else if (idx[a] == 0) {
            const Vq qp = vort_q(eos, ev_new, n, I + stride[a]);
            dub[a] = mul4(sub4(qp.ub, qn.ub), r1h[a]);
            duk[a] = mul4(sub4(qp.uk, qn.uk), r1h[a]);
            dT[a] = (qp.T - qn.T) * r1h[a];
        } else {
            const Vq qm = vort_q(eos, ev_new, n, I - stride[a]);
            dub[a] = mul4(sub4(qn.ub, qm.ub), r1h[a]);
            duk[a] = mul4(sub4(qn.uk, qm.uk), r1h[a]);
            dT[a] = (qn.T - qm.T) * r1h[a];
        }
    }
    const float rdt = p.r_dt, rtau = rcp(tau);
    // thermal vorticity (omega)
    const float4 dbt = mul4(sub4(qn.ub, qo.ub), rdt);
    const float4 dbx = dub[0], dby = dub[1], dbz = dub[2];
    const float hc = 0.5f * CV_HBARC;
    om[0 * n + I] = hc * (dby.w - dbz.z);
    om[1 * n + I] = hc * (dbz.y - dbx.w);
    om[2 * n + I] = hc * (dbx.z - dby.y);
    om[3 * n + I] = hc * ((dbt.w - dbz.x) + qn.ub.w * rtau);
    om[4 * n + I] = hc * (dby.x - dbt.z);
    om[5 * n + I] = hc * (dbt.y - dbx.x);
    // kinetic gradients with the eta Christoffel terms (omega_shear, omega_accT)
    const float4 dkt = mul4(sub4(qn.uk, qo.uk), rdt);
    const float4 dkx = duk[0], dky = duk[1];
    const float4 chr = make_float4(-qn.uk.w * rtau, 0.0f, 0.0f, -qn.uk.x * rtau);
    const float4 dkz = make_float4(duk[2].x + chr.x, duk[2].y + chr.y, duk[2].z + chr.z, duk[2].w + chr.w);
    const float4 rows[4] = {dkt, dkx, dky, dkz};
#pragma unroll
    for (int a = 0; a < 4; a++) {
        om[(6 + 4 * a + 0) * n + I] = 0.5f * rows[a].x;
        om[(6 + 4 * a + 1) * n + I] = 0.5f * rows[a].y;
        om[(6 + 4 * a + 2) * n + I] = 0.5f * rows[a].z;
        om[(6 + 4 * a + 3) * n + I] = 0.5f * rows[a].w;
    }
    // u^mu of the new field (umu4) and DU_nu = u^mu d_mu u_nu
    const float u0 = qn.uk.x, u1 = -qn.uk.y, u2 = -qn.uk.z, u3 = -qn.uk.w;
    float DU[4];
    const float c0[4] = {dkt.x, dkt.y, dkt.z, dkt.w}, c1[4] = {dkx.x, dkx.y, dkx.z, dkx.w};
    const float c2[4] = {dky.x, dky.y, dky.z, dky.w}, c3[4] = {dkz.x, dkz.y, dkz.z, dkz.w};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        DU[c] = u0 * c0[c] + u1 * c1[c] + u2 * c2[c] + u3 * c3[c];
        om[(22 + c) * n + I] = 0.5f * DU[c];
    }
    // omega_accT: temperature gradients
    const float T = qn.T, rT = rcp(T);
    const float dTt = (qn.T - qo.T) * rdt;
    const float gt = DU[0] - dTt * rT, gx = DU[1] - dT[0] * rT, gy = DU[2] - dT[1] * rT, gz = DU[3] - dT[2] * rT;
    const float k0 = qn.uk.x, k1 = qn.uk.y, k2 = qn.uk.z, k3 = qn.uk.w;
    om[26 * n + I] = 0.5f * (k2 * gz - k3 * gy);
    om[27 * n + I] = 0.5f * (-k1 * gz + k3 * gx);
    om[28 * n + I] = 0.5f * (k1 * gy - k2 * gx);
    om[29 * n + I] = 0.5f * (-k3 * gt + k0 * gz);
    om[30 * n + I] = 0.5f * (k2 * gt - k0 * gy);
    om[31 * n + I] = 0.5f * (-k1 * gt + k0 * gx);
}

// omega_mu (kernel_vorticity.cl:103-140): (T, omega^x, omega^y, omega^eta) with
// omega^mu = Omega^{mu nu} u_nu; zero where e < 0.1 efrz (compared in double as there).
// out: AoS rows of 4.
template <int EK>
__global__ __launch_bounds__(256) void k_vorticity_mu(PhysP p, EosP E, const float *__restrict__ ev,
                                                      const float *__restrict__ om, float efrz, float *out) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= p.n) return;
    EosT<EK> eos(E, E.tab);
    const int64_t n = p.n;
    const float e = ev[I], vx = ev[n + I], vy = ev[2 * n + I], vz = ev[3 * n + I];
    const float g = gamma3(vx, vy, vz);
    const float u0 = g * 1.0f, u1 = g * vx, u2 = g * vy, u3 = g * vz;
    const float tx = om[I], ty = om[n + I], tz = om[2 * n + I], xy = om[3 * n + I], xz = om[4 * n + I],
                yz = om[5 * n + I];
    float4 r = make_float4(eos.T(e), -u0 * tx - u1 * 0.0f - u2 * xy - u3 * xz,
                           -u0 * ty + u1 * xy - u2 * 0.0f - u3 * yz, -u0 * tz + u1 * xz + u2 * yz - u3 * 0.0f);
    if ((double)e < 0.1 * (double)efrz) r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    out[4 * I + 0] = r.x; out[4 * I + 1] = r.y; out[4 * I + 2] = r.z; out[4 * I + 3] = r.w;
}

}  // namespace cv
