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

// e(x, y) of each column (one thread per (i, j)), then e(x, y) H(eta) into the four planes
// of ev (velocities zero) one thread per cell, so that the field writes are coalesced
__global__ __launch_bounds__(256) void k_glauber_columns(PhysP p, cv_glauber g, float *edxy) {
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
    edxy[t] = kfac * gl_ed_transverse(g, ta, tb);
}

__global__ __launch_bounds__(256) void k_glauber_ini(PhysP p, cv_glauber g, const float *__restrict__ edxy, float *ev) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= p.n) return;
    const int64_t t = I / p.nz;
    const int k = (int)(I - t * p.nz);
    const float etas = (float)(k - p.nz / 2) * p.dz;
    const int64_t n = p.n;
    ev[I] = edxy[t] * gl_weight_eta(g, etas);
    ev[n + I] = 0.0f;
    ev[2 * n + I] = 0.0f;
    ev[3 * n + I] = 0.0f;
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

// ---------------------------------------------------------------------------------------
// Gaussian smearing of a parton list (kernel_gaussian_smearing_new.cl:11-120, used by
// SmearingP4X4 / CLVisc.smear_from_p4x4): every cell sums, over the partons closer than 10
// sigma, w exp(-d^2 / (2 SIGR^2)) (m_T cosh(Y - eta_s), p_x, p_y, m_T sinh(Y - eta_s)), with
// d the (x, y, eta_s) distance (the eta_s difference also scaled by SIGR, as there) and
// w = 1/(2 pi SIGR^2) sqrt(1/(2 pi SIGZ^2)) / TAU0; then KFACTOR, the fixed-point root solve
// of helper.h:112-125 for e, and (e, 0, 0, 0) into the field.  The per-parton quantities
// (eta_s, m_T, Y) do not depend on the cell and are computed once per parton.
struct SmearParton {
    float x, y, etas, mt, Y, px, py, pad;
};

__global__ __launch_bounds__(256) void k_smear_prep(const float *__restrict__ p4x4, int64_t np_, SmearParton *out) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= np_) return;
    const float *q = p4x4 + 8 * m;           // E px py pz t x y z
    const float E = q[0], px = q[1], py = q[2], pz = q[3], t = q[4], x = q[5], y = q[6], z = q[7];
    SmearParton s;
    s.x = x;
    s.y = y;
    s.etas = 0.5f * logf(rmax(t + z, CV_ACU)) - 0.5f * logf(rmax(t - z, CV_ACU));
    s.mt = sqrtf(rmax(E * E - pz * pz, px * px + py * py));
    s.Y = 0.5f * (logf(rmax(E + pz, CV_ACU)) - logf(rmax(E - pz, CV_ACU)));
    s.px = px;
    s.py = py;
    s.pad = 0.0f;
    out[m] = s;
}

// One workgroup per block of 16 x 8 x 8 cells (i, j, k; k fastest), each thread owning 4
// cells along i.  The partons pass through LDS in chunks of CV_SMEAR_CHUNK with what the
// block needs of each: squared distances to the block's 16 x, 8 y and 8 eta coordinates,
// and mt cosh(Y - eta), mt sinh(Y - eta) at its 8 eta (the same expressions as per pair,
// evaluated once per parton and block row instead of once per cell).  Every cell then sums
// its partons in list order with the reference's distance, cut and weight.
#define CV_SMEAR_CHUNK 128
#define CV_SMEAR_BI 16
#define CV_SMEAR_BJ 8
#define CV_SMEAR_BK 8
__global__ __launch_bounds__(256) void k_smear(PhysP p, EosP E, const SmearParton *__restrict__ par, int64_t np_,
                                               float one_o_2sigr2, float w, float kfactor, float *ev) {
    // rows padded by one word: the 8 rows a wavefront reads at one column fall in 8 banks
    __shared__ float4 s_dx[CV_SMEAR_CHUNK][CV_SMEAR_BI / 4];
    __shared__ float s_dy[CV_SMEAR_BJ][CV_SMEAR_CHUNK + 1], s_dz[CV_SMEAR_BK][CV_SMEAR_CHUNK + 1];
    __shared__ float s_ch[CV_SMEAR_BK][CV_SMEAR_CHUNK + 1], s_sh[CV_SMEAR_BK][CV_SMEAR_CHUNK + 1];
    __shared__ float s_px[CV_SMEAR_CHUNK], s_py[CV_SMEAR_CHUNK];
    const int tid = threadIdx.x;
    const int ap = tid / (CV_SMEAR_BJ * CV_SMEAR_BK);                 // 4 cells i0 + 4 ap + r
    const int b = (tid / CV_SMEAR_BK) % CV_SMEAR_BJ, c = tid % CV_SMEAR_BK;
    const int i0 = blockIdx.x * CV_SMEAR_BI, j0 = blockIdx.y * CV_SMEAR_BJ, k0 = blockIdx.z * CV_SMEAR_BK;
    const int j = j0 + b, k = k0 + c;
    float t0[4] = {0.0f, 0.0f, 0.0f, 0.0f}, t1[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float t2[4] = {0.0f, 0.0f, 0.0f, 0.0f}, t3[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int64_t n = 0; n < np_; n += CV_SMEAR_CHUNK) {
        const int cnt = (int)(np_ - n < CV_SMEAR_CHUNK ? np_ - n : CV_SMEAR_CHUNK);
        __syncthreads();
        // threads 0..127: x and y distances, px, py; threads 128..255: eta distances, cosh, sinh
        const int m = tid % CV_SMEAR_CHUNK;
        if (m < cnt) {
            const SmearParton q = par[n + m];
            if (tid < CV_SMEAR_CHUNK) {
#pragma unroll
                for (int u4 = 0; u4 < CV_SMEAR_BI / 4; u4++) {
                    float d[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const float d1 = (float)(i0 + 4 * u4 + r - p.nx / 2) * p.dx - q.x;
                        d[r] = d1 * d1;
                    }
                    s_dx[m][u4] = make_float4(d[0], d[1], d[2], d[3]);
                }
#pragma unroll
                for (int u = 0; u < CV_SMEAR_BJ; u++) {
                    const float d2 = (float)(j0 + u - p.ny / 2) * p.dy - q.y;
                    s_dy[u][m] = d2 * d2;
                }
                s_px[m] = q.px;
                s_py[m] = q.py;
            } else {
#pragma unroll
                for (int u = 0; u < CV_SMEAR_BK; u++) {
                    const float ge = (float)(k0 + u - p.nz / 2) * p.dz;
                    const float d3 = ge - q.etas;
                    s_dz[u][m] = d3 * d3;
                    s_ch[u][m] = q.mt * coshf(q.Y - ge);
                    s_sh[u][m] = q.mt * sinhf(q.Y - ge);
                }
            }
        }
        __syncthreads();
        for (int mm = 0; mm < cnt; mm++) {
            const float4 dxv = s_dx[mm][ap];
            const float sy = s_dy[b][mm], sz = s_dz[c][mm], chv = s_ch[c][mm], shv = s_sh[c][mm];
            const float px = s_px[mm], py = s_py[mm];
            const float dx[4] = {dxv.x, dxv.y, dxv.z, dxv.w};
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float dist = one_o_2sigr2 * (dx[r] + sy + sz);
                if (dist < 50.0f) {
                    const float delta = w * expf(-dist);
                    t0[r] = t0[r] + delta * chv;
                    t1[r] = t1[r] + delta * px;
                    t2[r] = t2[r] + delta * py;
                    t3[r] = t3[r] + delta * shv;
                }
            }
        }
    }
    if (j >= p.ny || k >= p.nz) return;
    Eos eos(E, E.tab);
    const int64_t n = p.n;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = i0 + 4 * ap + r;
        if (i >= p.nx) break;
        const float T0 = kfactor * t0[r], T1 = kfactor * t1[r], T2 = kfactor * t2[r], T3 = kfactor * t3[r];
        const float K2 = T1 * T1 + T2 * T2 + T3 * T3;
        const float K0 = T0 < CV_ACU ? CV_ACU : T0;
        // fixed-point root solve (helper.h:112-125)
        float E1 = K0, E0;
        int it = 0;
        for (;;) {
            E0 = E1;
            E1 = K0 - K2 / (K0 + eos.P(E1));
            it++;
            if (it > 100 || fabsf(E1 - E0) / rmax(fabsf(E1), CV_ACU) < CV_ACU) break;
        }
        const int64_t I = ((int64_t)i * p.ny + j) * p.nz + k;
        ev[I] = rmax(0.0f, E1);
        ev[n + I] = 0.0f;
        ev[2 * n + I] = 0.0f;
        ev[3 * n + I] = 0.0f;
    }
}

// change_longitudinal_profile (kernel_gaussian_smearing_new.cl:124-138): every cell of a
// column becomes (profile[k] e_mid, v_mid) of the column's mid-rapidity cell.  Two passes so
// that the mid-rapidity cell is read before it is rewritten (the reference's single pass
// races on it when profile[NZ/2] != 1).  mid: 4 planes of NX*NY.
__global__ __launch_bounds__(256) void k_long_profile_mid(PhysP p, const float *__restrict__ ev, float *mid) {
    const int64_t nxy = (int64_t)p.nx * p.ny;
    const int64_t ij = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= nxy) return;
    const int64_t I = ij * p.nz + p.nz / 2;
    for (int c = 0; c < 4; c++) mid[c * nxy + ij] = ev[c * p.n + I];
}
__global__ __launch_bounds__(256) void k_long_profile(PhysP p, const float *__restrict__ mid,
                                                      const float *__restrict__ prof, float *ev) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= p.n) return;
    const int64_t nxy = (int64_t)p.nx * p.ny;
    const int64_t ij = I / p.nz;
    ev[I] = prof[I - ij * p.nz] * mid[ij];
    for (int c = 1; c < 4; c++) ev[c * p.n + I] = mid[c * nxy + ij];
}

// Entropy density -> energy density of a field (from_sd_to_ed, pyvisc/ebe_trento.py:26-33:
// InterpolatedUnivariateSpline(eos.s, eos.ed, k=1), i.e. piecewise linear in s through the
// EOS nodes with linear extrapolation beyond the first and last segments), in double.
__global__ __launch_bounds__(256) void k_sd_to_ed(int64_t n, float *e, const double *__restrict__ s_nodes,
                                                  const double *__restrict__ e_nodes, int64_t m) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= n) return;
    const double s = e[I];
    int64_t lo;                              // the segment [lo, lo + 1] that s falls in (or the end one)
    if (s <= s_nodes[0]) {
        lo = 0;
    } else if (s >= s_nodes[m - 1]) {
        lo = m - 2;
    } else {
        lo = 0;
        int64_t up = m - 1;                  // s_nodes[lo] <= s < s_nodes[up]
        while (up - lo > 1) {
            const int64_t mid = (lo + up) >> 1;
            if (s_nodes[mid] <= s) lo = mid; else up = mid;
        }
    }
    const int64_t hi = lo + 1;
    const double t = (s - s_nodes[lo]) / (s_nodes[hi] - s_nodes[lo]);
    e[I] = (float)(e_nodes[lo] + t * (e_nodes[hi] - e_nodes[lo]));
}

}  // namespace cv
