// cv_kernels.h -- per-cell stage kernels of the KT + Israel-Stewart hydro step.
//
// One thread owns one cell and evaluates a whole reference stage for it:
//   k_ideal_stage : kt_src_christoffel + kt_src_along{x,y,z} + update_ev   (kernel_ideal.cl)
//   k_is_stage    : visc_src_christoffel + visc_src_along{x,y,z} + update_pimn (kernel_IS.cl)
//   k_kt_stage    : kt_src_christoffel + kt_src_along{x,y,z} + update_ev   (kernel_visc.cl)
//                   optionally followed by get_udiff (kernel_IS.cl:714-724)
// Fields are structure-of-arrays planes: component c of cell I sits at base[c*n + I].
// The 2-ghost constant extrapolation of the reference pencils (kernel_ideal.cl:57-62)
// becomes a clamp of the stencil index to the grid.
#pragma once
#include "cv_device.h"

#ifndef CV_PHASE_FENCE
#define CV_PHASE_FENCE() __builtin_amdgcn_sched_barrier(0)
#endif

namespace cv {

__device__ __forceinline__ float4 ld4(const float *__restrict__ b, int64_t n, int64_t I) {
    return make_float4(b[I], b[n + I], b[2 * n + I], b[3 * n + I]);
}
__device__ __forceinline__ void st4(float *b, int64_t n, int64_t I, float4 v) {
    b[I] = v.x; b[n + I] = v.y; b[2 * n + I] = v.z; b[3 * n + I] = v.w;
}
// Loads through a uniform plane base and a 32-bit byte offset (cells < 2^30 per field):
// lets the compiler use the scalar-base + 32-bit vector-offset form of global loads, so
// one offset register serves every plane of a field.
__device__ __forceinline__ float ldo(const float *__restrict__ base, uint32_t boff) {
    return *(const float *)((const char *)base + boff);
}
__device__ __forceinline__ float4 ld4o(const float *__restrict__ b, int64_t n, uint32_t boff) {
    return make_float4(ldo(b, boff), ldo(b + n, boff), ldo(b + 2 * n, boff), ldo(b + 3 * n, boff));
}
// Plane m of a field at byte offset boff of plane 0, as one 32-bit offset from the field's
// base (nb = bytes per plane): every field shares the uniform plane offsets m * nb, so the
// kernel keeps one base address per field instead of one per plane.
__device__ __forceinline__ float ldp(const float *__restrict__ base, uint32_t boff, uint32_t nb, int m) {
    return ldo(base, boff + (uint32_t)m * nb);
}
__device__ __forceinline__ float4 ld4p(const float *__restrict__ b, uint32_t boff, uint32_t nb) {
    return make_float4(ldp(b, boff, nb, 0), ldp(b, boff, nb, 1), ldp(b, boff, nb, 2), ldp(b, boff, nb, 3));
}
__device__ __forceinline__ void stp(float *base, uint32_t boff, uint32_t nb, int m, float v) {
    *(float *)((char *)base + (boff + (uint32_t)m * nb)) = v;
}
__device__ __forceinline__ float comp(float4 v, int c) {
    return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w));
}
__device__ __forceinline__ float4 half4(float4 a, float4 b) {
    return make_float4(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z), 0.5f * (a.w + b.w));
}

// index of the cell at offset s along axis a (clamped to the grid)
struct Cell {
    int64_t I;
    int i, j, k;
};
__device__ __forceinline__ Cell cell_of(const PhysP &p, int64_t I) {
    Cell c;
    c.I = I;
    c.i = (int)(I / p.sx);
    int64_t r = I - (int64_t)c.i * p.sx;
    c.j = (int)(r / p.nz);
    c.k = (int)(r - (int64_t)c.j * p.nz);
    return c;
}
template <int A>
__device__ __forceinline__ int64_t nb(const PhysP &p, const Cell &c, int s) {
    int t = A == 0 ? c.i : (A == 1 ? c.j : c.k);
    int L = A == 0 ? p.nx : (A == 1 ? p.ny : p.nz);
    int64_t st = A == 0 ? p.sx : (A == 1 ? p.sy : 1);
    int u = t + s;
    u = u < 0 ? 0 : (u > L - 1 ? L - 1 : u);
    return c.I + (int64_t)(u - t) * st;
}

template <int A>
__device__ __forceinline__ uint32_t nbo(const PhysP &p, const Cell &c, int s) {
    int t = A == 0 ? c.i : (A == 1 ? c.j : c.k);
    int L = A == 0 ? p.nx : (A == 1 ? p.ny : p.nz);
    uint32_t st = A == 0 ? (uint32_t)p.sx : (A == 1 ? (uint32_t)p.sy : 1u);
    int u = t + s;
    u = u < 0 ? 0 : (u > L - 1 ? L - 1 : u);
    return ((uint32_t)c.I + (uint32_t)(u - t) * st) * 4u;
}

// 10 components of pi^{mu nu}: index of (mu, nu) in the idx() order (helper.h:31)
__device__ __forceinline__ int IDX(int i, int j) {
    return (i < j) ? ((7 * i + 2 * j - i * i) / 2) : ((7 * j + 2 * i - j * j) / 2);
}
__constant__ const int PI_ROW[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};

// ---------------------------------------------------------------------------
// kt1d (helper.h:179-242): ideal KT flux difference for 5 stencil cells
// ---------------------------------------------------------------------------
template <int A>
__device__ __forceinline__ float4 kt1d_ideal(const PhysP &p, const Eos &eos, const float4 c[5]) {
    float pr_im1 = eos.P(c[1].x), pr_i = eos.P(c[2].x), pr_ip1 = eos.P(c[3].x), pr_ip2 = eos.P(c[4].x);
    float4 Tm1 = t0m(c[1], pr_im1), Ti = t0m(c[2], pr_i), Tp1 = t0m(c[3], pr_ip1), Tp2 = t0m(c[4], pr_ip2);
    float DA0[4], DA1[4], AL[4], AR[4], src[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        DA0[q] = slope(comp(Tm1, q), comp(Ti, q), comp(Tp1, q));
        DA1[q] = slope(comp(Ti, q), comp(Tp1, q), comp(Tp2, q));
        AL[q] = comp(Ti, q) + 0.5f * DA0[q];
        AR[q] = comp(Tp1, q) - 0.5f * DA1[q];
    }
    float pr_half = 0.5f * (pr_ip1 + pr_i);
    float vh = 0.5f * (comp(c[2], A + 1) + comp(c[3], A + 1));
    float lam = max_prop_speed(p, eos, half4(c[2], c[3]), vh);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float g0 = q == 0 ? 1.0f : 0.0f;
        float ga = q == A + 1 ? -1.0f : 0.0f;
        float Jp = (AR[q] + pr_half * g0) * vh - pr_half * ga;
        float Jm = (AL[q] + pr_half * g0) * vh - pr_half * ga;
        src[q] = 0.5f * (Jp + Jm) - 0.5f * lam * (AR[q] - AL[q]);
    }
    float pr_im2 = eos.P(c[0].x);
    float4 Tm2 = t0m(c[0], pr_im2);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        DA1[q] = DA0[q];
        DA0[q] = slope(comp(Tm2, q), comp(Tm1, q), comp(Ti, q));
        AL[q] = comp(Tm1, q) + 0.5f * DA0[q];
        AR[q] = comp(Ti, q) - 0.5f * DA1[q];
    }
    pr_half = 0.5f * (pr_im1 + pr_i);
    vh = 0.5f * (comp(c[1], A + 1) + comp(c[2], A + 1));
    lam = max_prop_speed(p, eos, half4(c[2], c[1]), vh);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float g0 = q == 0 ? 1.0f : 0.0f;
        float ga = q == A + 1 ? -1.0f : 0.0f;
        float Jp = (AR[q] + pr_half * g0) * vh - pr_half * ga;
        float Jm = (AL[q] + pr_half * g0) * vh - pr_half * ga;
        src[q] -= 0.5f * (Jp + Jm) - 0.5f * lam * (AR[q] - AL[q]);
    }
    return make_float4(src[0], src[1], src[2], src[3]);
}

template <int A>
__device__ __forceinline__ void ideal_axis(const PhysP &p, const Eos &eos, const StageP &st, const Cell &cl,
                                           const float *__restrict__ ev, float s[4]) {
    float4 c[5];
#pragma unroll
    for (int q = 0; q < 5; q++) c[q] = ld4(ev, p.n, nb<A>(p, cl, q - 2));
    float4 k = kt1d_ideal<A>(p, eos, c);
    float kk[4] = {k.x, k.y, k.z, k.w};
    float rdx = A == 0 ? p.r_dx : (A == 1 ? p.r_dy : st.r_taudz);
    bool riemann = p.flags & 1;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        if (riemann) s[m] = s[m] - kk[m] * (A == 0 ? p.r_dx : (A == 1 ? p.r_dy : p.r_dz));
        else s[m] = s[m] - (st.tau * kk[m]) * rdx;
    }
}

// jet_eloss_src (kernel_jet_eloss.cl:8-69), added to the stage source after the KT sweeps
// (ideal.py:240-246, visc.py:265-271): a Gaussian of width 0.3 fm around the jet, which moves
// from the start cell at the speed of light; dE/dx = 14 GeV/fm (T/0.3 GeV)^3 with T at the jet's
// cell of ev_st.  exp and the cube are taken in FP64 and rounded once (as the oracle does).
// Returns the four terms coef[q] * dE[q] that the caller adds to its source (by value and
// inlined: a private array or struct whose address reached a call would live in scratch
// memory in every launch of the kernel, jet or not).
__device__ __forceinline__ float4 jet_deposit(const PhysP &p, const Eos &eos, const JetP &j, float tau,
                                              const float *__restrict__ ev_st, const Cell &cl) {
    const float sig = 0.3f;
    const float dtau = tau - j.tau0;
    const float px = j.x0 + j.c * dtau, py = j.y0 + j.sn * dtau, pz = j.z0 + 0.0f * dtau;
    const int li = (int)(px * p.r_dx + (float)(p.nx / 2));
    const int lj = (int)(py * p.r_dy + (float)(p.ny / 2));
    const int lk = (int)(pz * p.r_dz + (float)(p.nz / 2));
    if (!(li >= 0 && li < p.nx && lj >= 0 && lj < p.ny && lk >= 0 && lk < p.nz))
        return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float x = (float)(cl.i - p.nx / 2) * p.dx;
    const float y = (float)(cl.j - p.ny / 2) * p.dy;
    const float z = (float)(cl.k - p.nz / 2) * p.dz;
    const float norm = (float)(1.0 / (double)(j.c3 * sig * sig * sig * tau));
    const float r2 = rcp(2.0f * sig * sig);
    const float g = (-(x - px)) * (x - px) * r2 - (y - py) * (y - py) * r2 - (z - pz) * (z - pz) * r2;
    const float f = norm * (float)exp((double)g);
    const float coef[4] = {tau * f, tau * f, tau * f, (tau * tau) * f};
    const float e = ev_st[(int64_t)li * p.sx + (int64_t)lj * p.sy + lk];
    const double tr = (double)(eos.T(e) * rcp(0.3f));
    const float dEdx = 14.0f * (float)(tr * tr * tr);
    const float dE[4] = {dEdx, dEdx * j.c, dEdx * j.sn, 0.0f};
    return make_float4(coef[0] * dE[0], coef[1] * dE[1], coef[2] * dE[2], coef[3] * dE[3]);
}
// s[q] + deposit[q], the order of the reference's d_Src + jet source
__device__ __forceinline__ void jet_add(float *s, float4 d) {
    s[0] = s[0] + d.x; s[1] = s[1] + d.y; s[2] = s[2] + d.z; s[3] = s[3] + d.w;
}

// ---------------------------------------------------------------------------
// CLIdeal.stepUpdate(step): christoffel + 3 sweeps + update_ev, one cell per thread
// ---------------------------------------------------------------------------
// cons_out (optional): write (T^tt, |T^ti|, T^t eta) of the root solve's input to three
// planes and change nothing else (the input of the learned seed).  seed_v (optional): start
// the root solve at this |v|.  iters_out (optional): iterations of the root solve.
__global__ __launch_bounds__(256) void k_ideal_stage(PhysP p, EosP E, StageP st,
                                                     const float *__restrict__ ev_st,
                                                     const float *ev1, float *Src, float *ev_out,
                                                     float *cons_out, const float *seed_v,
                                                     int *iters_out, JetP jet) {
    int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= p.n) return;
    Eos eos(E, E.tab);
    Cell cl = cell_of(p, I);
    const bool riemann = p.flags & 1;
    float s[4];
    // kt_src_christoffel (kernel_ideal.cl:3-36)
    if (st.step == 1) { s[0] = s[1] = s[2] = s[3] = 0.0f; }
    else { float4 t = ld4(Src, p.n, I); s[0] = t.x; s[1] = t.y; s[2] = t.z; s[3] = t.w; }
    {
        float4 e = ld4(ev_st, p.n, I);
        float pressure = eos.P(e.x);
        float u0 = gamma3(e.y, e.z, e.w);
        float Tzz = (e.x + pressure) * u0 * u0 * e.w * e.w + pressure;
        float Ttz = (e.x + pressure) * u0 * u0 * e.w;
        if (!riemann) { s[0] = s[0] - Tzz; s[1] = s[1] - 0.0f; s[2] = s[2] - 0.0f; s[3] = s[3] - Ttz; }
    }
    ideal_axis<0>(p, eos, st, cl, ev_st, s);
    ideal_axis<1>(p, eos, st, cl, ev_st, s);
    ideal_axis<2>(p, eos, st, cl, ev_st, s);
    if (jet.on & st.step) jet_add(s, jet_deposit(p, eos, jet, st.tau, ev_st, cl));
    if (!cons_out) st4(Src, p.n, I, make_float4(s[0], s[1], s[2], s[3]));
    // update_ev (kernel_ideal.cl:165-231)
    float4 e = ld4(ev1, p.n, I);
    float pressure = eos.P(e.x);
    float u0 = gamma3(e.y, e.z, e.w);
    float umu[4] = {u0 * 1.0f, u0 * e.y, u0 * e.z, u0 * e.w};
    float T0m[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float g0 = q == 0 ? 1.0f : 0.0f;
        float t = (e.x + pressure) * u0 * umu[q] - pressure * g0;
        T0m[q] = riemann ? t : t * st.old_time;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) T0m[q] = T0m[q] + (s[q] * p.dt) * st.r_step;
    float T00, T01, T02, T03;
    if (riemann) {
        T00 = rmax(CV_ACU, T0m[0]);
        T01 = fabsf(T0m[1]) < CV_ACU ? 0.0f : T0m[1];
        T02 = fabsf(T0m[2]) < CV_ACU ? 0.0f : T0m[2];
        T03 = fabsf(T0m[3]) < CV_ACU ? 0.0f : T0m[3];
    } else {
        T00 = rmax(CV_ACU, T0m[0]) * st.r_tau;
        T01 = fabsf(T0m[1]) < CV_ACU ? 0.0f : T0m[1] * st.r_tau;
        T02 = fabsf(T0m[2]) < CV_ACU ? 0.0f : T0m[2] * st.r_tau;
        T03 = fabsf(T0m[3]) < CV_ACU ? 0.0f : T0m[3] * st.r_tau;
    }
    float M = sqrtf(T01 * T01 + T02 * T02 + T03 * T03);
    const float SC = 0.999f;
    if (M > T00) {
        float sc = (SC * T00) * rcp(M);
        T01 *= sc; T02 *= sc; T03 *= sc;
        M = SC * T00;
    }
    if (cons_out) {
        cons_out[I] = T00; cons_out[p.n + I] = M; cons_out[2 * p.n + I] = T03;
        return;
    }
    float ed_find;
    int it = 0;
    if ((p.flags & 8) && M > p.fp64_vmin * T00)
        ed_find = (float)root_newton_d(E, E.tab, (double)T00, (double)M, 0.0);
    else if (seed_v)
        ed_find = root_newton_from(eos, T00, M, 0.0f, false, fminf(fmaxf(seed_v[I], 0.0f), 0.9999f), &it);
    else
        ed_find = root_newton_from(eos, T00, M, 0.0f, false, 0.5f * (0.0f + 1.0f), &it);
    if (iters_out) iters_out[I] = it;
    ed_find = rmax(0.0f, ed_find);
    float pr = eos.P(ed_find);
    float repv = rcp(rmax(CV_ACU, T00 + pr));
    st4(ev_out, p.n, I, make_float4(ed_find, T01 * repv, T02 * repv, T03 * repv));
}

// Root solve of given conserved states (T^tt, |T^ti|) from the middle of [0, 1] or from a
// seed |v|: e, |v| = M / (T^tt + P(e)) and the iteration count (evaluation of the seed).
__global__ __launch_bounds__(256) void k_root_eval(EosP E, const float *__restrict__ T00, const float *__restrict__ M,
                                                   const float *__restrict__ seed_v, int64_t n, float *e_out,
                                                   float *v_out, int *iters) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    Eos eos(E, E.tab);
    const float t = T00[q], m = M[q];
    int it = 0;
    const float v0 = seed_v ? fminf(fmaxf(seed_v[q], 0.0f), 0.9999f) : 0.5f * (0.0f + 1.0f);
    float e = rmax(0.0f, root_newton_from<false>(eos, t, m, 0.0f, false, v0, &it));
    e_out[q] = e;
    v_out[q] = m * rcp(rmax(CV_ACU, t + eos.P(e)));
    iters[q] = it;
}

// ---------------------------------------------------------------------------
// kt1d_real (helper.h:249-285)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float kt1d_real(float Q_im2, float Q_im1, float Q_i, float Q_ip1, float Q_ip2,
                                           float v_mh, float v_ph, float lam_mh, float lam_ph) {
    float DA0 = slope(Q_im1, Q_i, Q_ip1);
    float DA1 = slope(Q_i, Q_ip1, Q_ip2);
    float AL = Q_i + 0.5f * DA0;
    float AR = Q_ip1 - 0.5f * DA1;
    float Jp = AR * v_ph, Jm = AL * v_ph;
    float src = 0.5f * (Jp + Jm) - 0.5f * lam_ph * (AR - AL);
    DA1 = DA0;
    DA0 = slope(Q_im2, Q_im1, Q_i);
    AL = Q_im1 + 0.5f * DA0;
    AR = Q_i - 0.5f * DA1;
    Jp = AR * v_mh; Jm = AL * v_mh;
    src -= 0.5f * (Jp + Jm) - 0.5f * lam_mh * (AR - AL);
    return src;
}

// visc_src_along{x,y,z} for one cell (kernel_IS.cl:74-318)
template <int A, bool BULK>
__device__ __forceinline__ void is_axis(const PhysP &p, const Eos &eos, const StageP &st, const Cell &cl,
                                        const float *__restrict__ ev, const float *__restrict__ pi,
                                        const float *__restrict__ bk, float s[10], float &bs, float ud[4]) {
    int64_t J[5];
    float4 e[5];
    float u0[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        J[q] = nb<A>(p, cl, q - 2);
        e[q] = ld4(ev, p.n, J[q]);
        u0[q] = gamma3(e[q].y, e[q].z, e[q].w);
    }
    float v_mh = 0.5f * (comp(e[1], A + 1) + comp(e[2], A + 1));
    float v_ph = 0.5f * (comp(e[3], A + 1) + comp(e[2], A + 1));
    float lam_mh = max_prop_speed(p, eos, half4(e[1], e[2]), v_mh);
    float lam_ph = max_prop_speed(p, eos, half4(e[3], e[2]), v_ph);
    float um1[4] = {u0[1] * 1.0f, u0[1] * e[1].y, u0[1] * e[1].z, u0[1] * e[1].w};
    float ui[4] = {u0[2] * 1.0f, u0[2] * e[2].y, u0[2] * e[2].z, u0[2] * e[2].w};
    float up1[4] = {u0[3] * 1.0f, u0[3] * e[3].y, u0[3] * e[3].z, u0[3] * e[3].w};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        float d = slope(um1[c], ui[c], up1[c]);
        if (A == 0) ud[c] = d * p.r_dx;
        else if (A == 1) ud[c] = d * p.r_dy;
        else {
            float chr = ((c == 0 ? e[2].w : (c == 3 ? 1.0f : 0.0f)) * u0[2]) * st.r_tau;
            ud[c] = d * st.r_taudz + chr;
        }
    }
    float rdx = A == 0 ? p.r_dx : (A == 1 ? p.r_dy : st.r_taudz);
#pragma unroll
    for (int mn = 0; mn < 10; mn++) {
        const float *pl = pi + (int64_t)mn * p.n;
        float k = kt1d_real(u0[0] * pl[J[0]], u0[1] * pl[J[1]], u0[2] * pl[J[2]], u0[3] * pl[J[3]],
                            u0[4] * pl[J[4]], v_mh, v_ph, lam_mh, lam_ph);
        s[mn] = s[mn] - k * rdx;
    }
    if (BULK) {
        float k = kt1d_real(u0[0] * bk[J[0]], u0[1] * bk[J[1]], u0[2] * bk[J[2]], u0[3] * bk[J[3]],
                            u0[4] * bk[J[4]], v_mh, v_ph, lam_mh, lam_ph);
        bs = bs - k * rdx;
    }
}

// PiPi self coupling (GUBSER_VISC_TEST, kernel_IS.cl:322-347)
__device__ __forceinline__ float dot4(const float *a, const float *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}
__device__ float pipi(int lam, int mu, int nu, const float *pim, const float *u) {
    const float g[4] = {1.0f, -1.0f, -1.0f, -1.0f};
    float A[4] = {pim[IDX(0, lam)], -pim[IDX(1, lam)], -pim[IDX(2, lam)], -pim[IDX(3, lam)]};
    float D[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) D[a][b] = (a == b ? g[a] : 0.0f) - u[a] * u[b];
    float first = dot4(D[mu], A) * dot4(D[nu], A);
    float temp[4] = {dot4(D[0], A), dot4(D[1], A), dot4(D[2], A), dot4(D[3], A)};
    float second = -1.0f / 3.0f * ((mu == nu ? g[mu] : 0.0f) - u[mu] * u[nu]) * dot4(temp, A);
    return first + second;
}

// Vorticity couplings (PIMUNU_OMEGA_COUPLING, OMEGA_OMEGA_COUPLING, kernel_IS.cl:349-470):
// Omega^{mu nu} = 1/2 Delta^mu_a Delta^nu_b omega^{ab} and the projected products
// pi^<mu_lam Omega^nu>lam, Omega^<mu_lam Omega^nu>lam, in the reference's operand order.
__device__ float omega_proj(int mu, int nu, const float *om, const float *u) {
    const float ul[4] = {u[0], -u[1], -u[2], -u[3]};
    float D[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) D[a][b] = (a == b ? 1.0f : 0.0f) - u[a] * ul[b];
    const float A[4][4] = {{0.0f, om[0], om[1], om[2]}, {-om[0], 0.0f, om[3], om[4]},
                           {-om[1], -om[3], 0.0f, om[5]}, {-om[2], -om[4], -om[5], 0.0f}};
    float temp[4] = {dot4(D[nu], A[0]), dot4(D[nu], A[1]), dot4(D[nu], A[2]), dot4(D[nu], A[3])};
    return 0.5f * dot4(D[mu], temp);
}
__device__ __forceinline__ void omega_row(int lam, const float *Om, float B[4]) {
    if (lam == 0) { B[0] = 0.0f; B[1] = Om[0]; B[2] = Om[1]; B[3] = Om[2]; }
    else if (lam == 1) { B[0] = Om[0]; B[1] = 0.0f; B[2] = Om[3]; B[3] = Om[4]; }
    else if (lam == 2) { B[0] = Om[1]; B[1] = -Om[3]; B[2] = 0.0f; B[3] = Om[5]; }
    else { B[0] = Om[2]; B[1] = -Om[4]; B[2] = -Om[5]; B[3] = 0.0f; }
}
__device__ float pi_omega(int lam, int mu, int nu, const float *pim, const float *Om, const float *u) {
    const float g[4] = {1.0f, -1.0f, -1.0f, -1.0f};
    float A[4] = {pim[IDX(0, lam)], -pim[IDX(1, lam)], -pim[IDX(2, lam)], -pim[IDX(3, lam)]};
    float B[4];
    omega_row(lam, Om, B);
    float D[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) D[a][b] = (a == b ? g[a] : 0.0f) - u[a] * u[b];
    float first = 0.5f * (dot4(D[mu], A) * dot4(D[nu], B) + dot4(D[mu], B) * dot4(D[nu], A));
    float temp[4] = {dot4(D[0], A), dot4(D[1], A), dot4(D[2], A), dot4(D[3], A)};
    float second = -1.0f / 3.0f * ((mu == nu ? g[mu] : 0.0f) - u[mu] * u[nu]) * dot4(temp, B);
    return first + second;
}
__device__ float omega_omega(int lam, int mu, int nu, const float *Om, const float *u) {
    const float g[4] = {1.0f, -1.0f, -1.0f, -1.0f};
    float B[4];
    omega_row(lam, Om, B);
    float D[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) D[a][b] = (a == b ? g[a] : 0.0f) - u[a] * u[b];
    float first = dot4(D[mu], B) * dot4(D[nu], B);
    float temp[4] = {dot4(D[0], B), dot4(D[1], B), dot4(D[2], B), dot4(D[3], B)};
    float second = -1.0f / 3.0f * ((mu == nu ? g[mu] : 0.0f) - u[mu] * u[nu]) * dot4(temp, B);
    return first + second;
}
// the six projected components (Omega^{01}, 02, 03, 12, 13, 23) from the derivatives of u
__device__ void omega_cell(const float *udt, const float *udx, const float *udy, const float *udz, const float *u,
                           float Om[6]) {
    const float om[6] = {udt[1] - udx[0], udt[2] - udy[0], udt[3] - udz[0],
                         udx[2] - udy[1], udx[3] - udz[1], udy[3] - udz[2]};
    Om[0] = omega_proj(0, 1, om, u); Om[1] = omega_proj(0, 2, om, u); Om[2] = omega_proj(0, 3, om, u);
    Om[3] = omega_proj(1, 2, om, u); Om[4] = omega_proj(1, 3, om, u); Om[5] = omega_proj(2, 3, om, u);
}
// the coupling terms of update_pimn's source for component (mu, nu) (kernel_IS.cl:636-683)
__device__ __forceinline__ float couple_src(float src, int flags, int mu, int nu, const float *pim, const float *Om,
                                            const float *u, float one_over_taupi) {
    if (flags & CV_FLAG_PIMN_OMEGA)
        src += 2.0f * (pi_omega(0, mu, nu, pim, Om, u) - pi_omega(1, mu, nu, pim, Om, u)
                       - pi_omega(2, mu, nu, pim, Om, u) - pi_omega(3, mu, nu, pim, Om, u));
    if (flags & CV_FLAG_OMEGA_OMEGA)
        src -= 1.0f * one_over_taupi * (omega_omega(0, mu, nu, Om, u) - omega_omega(1, mu, nu, Om, u)
                                        - omega_omega(2, mu, nu, Om, u) - omega_omega(3, mu, nu, Om, u));
    return src;
}

// ---------------------------------------------------------------------------
// CLVisc.IS_stepUpdate(step): one cell per thread
//   ev_st/pi_st/bk_st : d_ev[step], d_pi[step] (stencil)
//   ev1, ev2          : d_ev[1], d_ev[2]; pi1, bk1 : start point d_pi[1]
//   pi_out, bk_out    : d_pi[3-step] (may alias pi1 at step 2: read-before-write per cell)
// ---------------------------------------------------------------------------
template <bool BULK>
__global__ __launch_bounds__(256) void k_is_stage(PhysP p, EosP E, StageP st,
        const float *__restrict__ ev_st, const float *__restrict__ pi_st, const float *__restrict__ bk_st,
        const float *ev1, const float *ev2, const float *pi1, const float *bk1,
        const float *__restrict__ udiff, float *IS_src, float *bsrc, float *pi_out, float *bk_out, float *good) {
    int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= p.n) return;
    Eos eos(E, E.tab);
    Cell cl = cell_of(p, I);
    const int step = st.step;
    float s[10];
    float bs = 0.0f;
    if (step == 1) {
#pragma unroll
        for (int m = 0; m < 10; m++) s[m] = 0.0f;
    } else {
#pragma unroll
        for (int m = 0; m < 10; m++) s[m] = IS_src[(int64_t)m * p.n + I];
        if (BULK) bs = bsrc[I];
    }
    // visc_src_christoffel (kernel_IS.cl:41-67)
    {
        float4 e = ld4(ev_st, p.n, I);
        float uz = e.w * gamma3(e.y, e.z, e.w);
        float q[10];
#pragma unroll
        for (int m = 0; m < 10; m++) q[m] = pi_st[(int64_t)m * p.n + I];
        s[0] -= (2.0f * uz * q[3]) * st.r_tau;
        s[1] -= (uz * q[6]) * st.r_tau;
        s[2] -= (uz * q[8]) * st.r_tau;
        s[3] -= (uz * (q[0] + q[9])) * st.r_tau;
        s[6] -= (uz * q[1]) * st.r_tau;
        s[8] -= (uz * q[2]) * st.r_tau;
        s[9] -= (2.0f * uz * q[3]) * st.r_tau;
    }
    float udx[4], udy[4], udz[4];
    is_axis<0, BULK>(p, eos, st, cl, ev_st, pi_st, bk_st, s, bs, udx);
    is_axis<1, BULK>(p, eos, st, cl, ev_st, pi_st, bk_st, s, bs, udy);
    is_axis<2, BULK>(p, eos, st, cl, ev_st, pi_st, bk_st, s, bs, udz);

    // update_pimn (kernel_IS.cl:501-709)
    float4 e1 = ld4(ev1, p.n, I), e2 = ld4(ev2, p.n, I);
    float g1 = gamma3(e1.y, e1.z, e1.w), g2 = gamma3(e2.y, e2.z, e2.w);
    float u_old[4] = {g1 * 1.0f, g1 * e1.y, g1 * e1.z, g1 * e1.w};
    float u_new[4] = {g2 * 1.0f, g2 * e2.y, g2 * e2.z, g2 * e2.w};
    float udt[4];
#pragma unroll
    for (int c = 0; c < 4; c++) udt[c] = (u_new[c] - u_old[c]) * p.r_dt;
    if (step == 1) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            float ud_ = udiff[(int64_t)c * p.n + I];
            udt[c] = ud_ * p.r_dt;
            u_new[c] = u_old[c] + ud_;
        }
    }
    float da[4][4];
#pragma unroll
    for (int c = 0; c < 4; c++) { da[c][0] = udt[c]; da[c][1] = udx[c]; da[c][2] = udy[c]; da[c][3] = udz[c]; }
    float u[4], DU[4];
#pragma unroll
    for (int c = 0; c < 4; c++) u[c] = step == 1 ? u_old[c] : u_new[c];
#pragma unroll
    for (int c = 0; c < 4; c++) DU[c] = dot4(u, da[c]);
    float ed_step = step == 1 ? e1.x : e2.x;
    float theta = udt[0] + udx[1] + udy[2] + udz[3];
    float4 cpTs = eos.raw(ed_step);
    float temperature = cpTs.z;
    float local_etaos = etaos(p, temperature);
    float etav = local_etaos * cpTs.w * CV_HBARC;
    float one_over_taupi = temperature * rcp(3.0f * rmax(CV_ACU, local_etaos) * CV_HBARC);
    float coef_pipi = 0.0f;
    const bool gubser = p.flags & 2;
    if (gubser) {
        float q4 = sqrtf(sqrtf(ed_step));
        float ETAOS = p.etaos_ymin;
        float etavH = (ETAOS * 4.0f) * p.r_3;
        float taupiH = (p.lam1 * p.lam1 * etavH) * p.r_3;
        one_over_taupi = rcp(taupiH * rcp(q4));
        coef_pipi = p.lam1 * rcp(ed_step);
        etav = ((ETAOS * 4.0f) * p.r_3) * (sqrtf(ed_step) * q4);
    }
    float pi2[10];
#pragma unroll
    for (int m = 0; m < 10; m++) pi2[m] = pi_st[(int64_t)m * p.n + I];
    const bool cpl = (p.flags & (CV_FLAG_PIMN_OMEGA | CV_FLAG_OMEGA_OMEGA)) != 0;
    float Om[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (cpl) omega_cell(udt, udx, udy, udz, u, Om);
    const float gd[4] = {1.0f, -1.0f, -1.0f, -1.0f};
    float max_abs = 0.0f;
    float out[10];
    float rden = rcp(u_new[0] + p.dt * one_over_taupi);
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
#pragma unroll
        for (int nu = mu; nu < 4; nu++) {
            float gmunu = mu == nu ? gd[mu] : 0.0f;
            float sigma = gd[mu] * da[nu][mu] + gd[nu] * da[mu][nu] - (u[mu] * DU[nu] + u[nu] * DU[mu])
                          - 2.0f / 3.0f * (gmunu - u[mu] * u[nu]) * theta;
            int mn = IDX(mu, nu);
            if (u[0] > 10000.0f) sigma = 0.0f;
            float piNS = etav * sigma;
            float pi_old = pi1[(int64_t)mn * p.n + I] * u_old[0];
            float src = (-pi2[mn] * theta) * p.r_3 - (pi2[mn] * u[0]) * st.r_tau;
#pragma unroll
            for (int l = 0; l < 4; l++)
                src -= (u[mu] * pi2[IDX(nu, l)] + u[nu] * pi2[IDX(mu, l)]) * DU[l] * gd[l];
            if (gubser)
                src -= one_over_taupi * coef_pipi * (pipi(0, mu, nu, pi2, u) - pipi(1, mu, nu, pi2, u)
                                                     - pipi(2, mu, nu, pi2, u) - pipi(3, mu, nu, pi2, u));
            if (cpl) src = couple_src(src, p.flags, mu, nu, pi2, Om, u, one_over_taupi);
            s[mn] += src;
            pi_old = (pi_old + (s[mn] * p.dt) * st.r_step + p.dt * one_over_taupi * piNS) * rden;
            if (fabsf(pi_old) > max_abs) max_abs = fabsf(pi_old);
            out[mn] = pi_old;
        }
    }
    float bout = 0.0f;
    if (BULK) {
        float zs = zetaos(p, temperature);
        float zetav = zs * cpTs.w * CV_HBARC;
        float oot_b = temperature * rcp(3.0f * rmax(CV_ACU, zs) * CV_HBARC);
        float b2 = bk_st[I];
        float bsrc_ = (b2 * theta) * p.r_3 - (b2 * u[0]) * st.r_tau;
        bs += bsrc_;
        float b_old = bk1[I] * u_old[0];
        bout = (b_old + (bs * p.dt) * st.r_step - p.dt * oot_b * zetav * theta) * rcp(u_new[0] + p.dt * oot_b);
        bout = bulk_regulate(bout, eos.P(e2.x), e2.x);
        if (fabsf(bout) > max_abs) max_abs = fabsf(bout);
    }
    float pre = eos.P(e2.x);
    enthalpy_cap<BULK>(p.zmax, out, bout, max_abs, e2.x, pre);
    float T00 = (e2.x + pre) * u_new[0] * u_new[0] - pre;
    float gc = 1.0f;
    if (max_abs > T00) {
#pragma unroll
        for (int m = 0; m < 10; m++) out[m] = 0.0f;
        bout = 0.0f;
        gc = 0.0f;
    }
#pragma unroll
    for (int m = 0; m < 10; m++) {
        IS_src[(int64_t)m * p.n + I] = s[m];
        pi_out[(int64_t)m * p.n + I] = out[m];
    }
    if (BULK) { bsrc[I] = bs; bk_out[I] = bout; }
    good[I] = gc;
}

// ---------------------------------------------------------------------------
// kt1d_visc (kernel_visc.cl:60-127) for one cell and axis
// ---------------------------------------------------------------------------
template <int A, bool BULK>
__device__ __forceinline__ void kt_axis(const PhysP &p, const Eos &eos, const StageP &st, const Cell &cl,
                                        const float *__restrict__ ev, const float *__restrict__ pi,
                                        const float *__restrict__ bk, float s[4]) {
    float4 e[5];
    float pr[5], Q[5][4], pim0[5][4], pimi[5][4];
    const float tau = st.tau;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        int64_t J = nb<A>(p, cl, i - 2);
        e[i] = ld4(ev, p.n, J);
        pr[i] = eos.P(e[i].x) + (BULK ? bk[J] : 0.0f);
        float4 T = t0m(e[i], pr[i]);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            pim0[i][q] = pi[(int64_t)PI_ROW[0][q] * p.n + J];
            pimi[i][q] = pi[(int64_t)PI_ROW[A + 1][q] * p.n + J];
            Q[i][q] = tau * (comp(T, q) + pim0[i][q]);
        }
    }
    float vip = 0.5f * (comp(e[2], A + 1) + comp(e[3], A + 1));
    float vim = 0.5f * (comp(e[1], A + 1) + comp(e[2], A + 1));
    float DA0[4], DA1[4], AL[4], AR[4], src[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        DA0[q] = slope(Q[1][q], Q[2][q], Q[3][q]);
        DA1[q] = slope(Q[2][q], Q[3][q], Q[4][q]);
        AL[q] = Q[2][q] + 0.5f * DA0[q];
        AR[q] = Q[3][q] - 0.5f * DA1[q];
    }
    float pr_half = 0.5f * (pr[2] + pr[3]);
    float lam = max_prop_speed(p, eos, half4(e[2], e[3]), vip);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float g0 = q == 0 ? 1.0f : 0.0f;
        float ga = q == A + 1 ? -1.0f : 0.0f;
        float p0h = 0.5f * (pim0[2][q] + pim0[3][q]);
        float pih = 0.5f * (pimi[2][q] + pimi[3][q]);
        float Jp = (AR[q] + pr_half * tau * g0 - tau * p0h) * vip - pr_half * tau * ga + tau * pih;
        float Jm = (AL[q] + pr_half * tau * g0 - tau * p0h) * vip - pr_half * tau * ga + tau * pih;
        src[q] = 0.5f * (Jp + Jm) - 0.5f * lam * (AR[q] - AL[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        DA1[q] = DA0[q];
        DA0[q] = slope(Q[0][q], Q[1][q], Q[2][q]);
        AL[q] = Q[1][q] + 0.5f * DA0[q];
        AR[q] = Q[2][q] - 0.5f * DA1[q];
    }
    pr_half = 0.5f * (pr[1] + pr[2]);
    lam = max_prop_speed(p, eos, half4(e[2], e[1]), vim);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float g0 = q == 0 ? 1.0f : 0.0f;
        float ga = q == A + 1 ? -1.0f : 0.0f;
        float p0h = 0.5f * (pim0[1][q] + pim0[2][q]);
        float pih = 0.5f * (pimi[1][q] + pimi[2][q]);
        float Jp = (AR[q] + pr_half * tau * g0 - tau * p0h) * vim - pr_half * tau * ga + tau * pih;
        float Jm = (AL[q] + pr_half * tau * g0 - tau * p0h) * vim - pr_half * tau * ga + tau * pih;
        src[q] -= 0.5f * (Jp + Jm) - 0.5f * lam * (AR[q] - AL[q]);
    }
    float rdx = A == 0 ? p.r_dx : (A == 1 ? p.r_dy : st.r_taudz);
#pragma unroll
    for (int m = 0; m < 4; m++) s[m] = s[m] - src[m] * rdx;
}

// ---------------------------------------------------------------------------
// CLVisc.visc_stepUpdate(step): one cell per thread
//   ev_st/pi_st/bk_st : d_ev[step], d_pi[step] (stencil)
//   ev1               : d_ev[1] start point; pim0/bk0 : d_pi[0]; pinew/bknew : d_pi[3-step]
//   ev_out            : d_ev[3-step] (may alias ev1 at step 2)
//   udiff             : when non-null, receives u(ev_out) - u(ev1) (get_udiff)
// ---------------------------------------------------------------------------
template <bool BULK>
__global__ __launch_bounds__(256) void k_kt_stage(PhysP p, EosP E, StageP st,
        const float *__restrict__ ev_st, const float *__restrict__ pi_st, const float *__restrict__ bk_st,
        const float *ev1, const float *pim0, const float *bk0, const float *pinew, const float *bknew,
        float *Src, float *ev_out, float *udiff, JetP jet) {
    int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= p.n) return;
    Eos eos(E, E.tab);
    Cell cl = cell_of(p, I);
    const bool riemann = p.flags & 1;
    float s[4];
    if (st.step == 1) { s[0] = s[1] = s[2] = s[3] = 0.0f; }
    else { float4 t = ld4(Src, p.n, I); s[0] = t.x; s[1] = t.y; s[2] = t.z; s[3] = t.w; }
    // kt_src_christoffel (kernel_visc.cl:11-53)
    {
        float4 e = ld4(ev_st, p.n, I);
        float u0 = gamma3(e.y, e.z, e.w);
        float bulkpi = BULK ? bk_st[I] : 0.0f;
        float pressure = eos.P(e.x) + bulkpi;
        float Tzz = (e.x + pressure) * u0 * u0 * e.w * e.w + pressure + pi_st[9 * p.n + I];
        float Ttz = (e.x + pressure) * u0 * u0 * e.w + pi_st[3 * p.n + I];
        s[0] = s[0] - Tzz; s[1] = s[1] - 0.0f; s[2] = s[2] - 0.0f; s[3] = s[3] - Ttz;
        if (riemann) {
            float Ttt = (e.x + pressure) * u0 * u0 - pressure;
            float Ttx = (e.x + pressure) * u0 * u0 * e.y;
            float Tty = (e.x + pressure) * u0 * u0 * e.z;
            s[0] = s[0] + (Tzz + Ttt); s[1] = s[1] + Ttx; s[2] = s[2] + Tty; s[3] = s[3] + 3.0f * Ttz;
        }
    }
    kt_axis<0, BULK>(p, eos, st, cl, ev_st, pi_st, bk_st, s);
    kt_axis<1, BULK>(p, eos, st, cl, ev_st, pi_st, bk_st, s);
    kt_axis<2, BULK>(p, eos, st, cl, ev_st, pi_st, bk_st, s);
    if (jet.on & st.step) jet_add(s, jet_deposit(p, eos, jet, st.tau, ev_st, cl));
    st4(Src, p.n, I, make_float4(s[0], s[1], s[2], s[3]));
    // update_ev (kernel_visc.cl:322-393)
    float4 e = ld4(ev1, p.n, I);
    float pressure = eos.P(e.x) + (BULK ? bk0[I] : 0.0f);
    float u0 = gamma3(e.y, e.z, e.w);
    float umu[4] = {u0 * 1.0f, u0 * e.y, u0 * e.z, u0 * e.w};
    float T0m[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float g0 = q == 0 ? 1.0f : 0.0f;
        T0m[q] = ((e.x + pressure) * u0 * umu[q] - pressure * g0 + pim0[(int64_t)q * p.n + I]) * st.old_time;
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
        T0m[q] = T0m[q] + (s[q] * p.dt) * st.r_step - pinew[(int64_t)q * p.n + I] * st.tau;
    float T00 = rmax(CV_ACU, T0m[0]) * st.r_tau;
    float T01 = fabsf(T0m[1]) < CV_ACU ? 0.0f : T0m[1] * st.r_tau;
    float T02 = fabsf(T0m[2]) < CV_ACU ? 0.0f : T0m[2] * st.r_tau;
    float T03 = fabsf(T0m[3]) < CV_ACU ? 0.0f : T0m[3] * st.r_tau;
    float M = sqrtf(T01 * T01 + T02 * T02 + T03 * T03);
    const float SC = 0.999f;
    if (M > T00) {
        float sc = (SC * T00) * rcp(M);
        T01 *= sc; T02 *= sc; T03 *= sc;
        M = SC * T00;
    }
    float bnew = BULK ? bknew[I] : 0.0f;
    float ed_find;
    if ((p.flags & 8) && M > p.fp64_vmin * T00)
        ed_find = (float)root_newton_d(E, E.tab, (double)T00, (double)M, (double)bnew, BULK);
    else
        ed_find = root_newton(eos, T00, M, bnew, BULK);
    if (ed_find < CV_ACU) { T01 = 0.0f; T02 = 0.0f; T03 = 0.0f; }
    ed_find = rmax(0.0f, ed_find);
    float pr = eos.P(ed_find) + bnew;
    if (BULK) pr = rmax(pr, 0.0f);
    float repv = rcp(rmax(CV_ACU, T00 + pr));
    float4 en = make_float4(ed_find, T01 * repv, T02 * repv, T03 * repv);
    st4(ev_out, p.n, I, en);
    if (udiff) {
        float gn = gamma3(en.y, en.z, en.w);
        float un[4] = {gn * 1.0f, gn * en.y, gn * en.z, gn * en.w};
        float uo[4] = {u0 * 1.0f, u0 * e.y, u0 * e.z, u0 * e.w};
#pragma unroll
        for (int c = 0; c < 4; c++) udiff[(int64_t)c * p.n + I] = un[c] - uo[c];
    }
}

// ---------------------------------------------------------------------------
// Fused stage kernel of the step path: IS_stepUpdate(step) and visc_stepUpdate(step) of
// one cell in one thread.  The stencil of each axis is loaded once and serves the
// Israel-Stewart flux, the KT flux and the velocity gradient; lambda of each face is
// evaluated once for both (the reference evaluates the same expression twice).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float4 t0m_u(float4 ev, float pr, float u0) {
    float a = (ev.x + pr) * u0;
    return make_float4(a * (1.0f * u0) - 1.0f * pr, a * (ev.y * u0) - 0.0f * pr,
                       a * (ev.z * u0) - 0.0f * pr, a * (ev.w * u0) - 0.0f * pr);
}

template <int A, bool BULK>
__device__ __forceinline__ void visc_axis(const PhysP &p, const Eos &eos, const StageP &st, const Cell &cl,
                                          const float *__restrict__ ev, const float *__restrict__ pi,
                                          const float *__restrict__ bk, float s_is[10], float &bs,
                                          float s_kt[4], float ud[4]) {
    uint32_t J[5];
    float4 e[5];
    float u0[5], pr[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        J[q] = nbo<A>(p, cl, q - 2);
        e[q] = ld4o(ev, p.n, J[q]);
        u0[q] = gamma3(e[q].y, e[q].z, e[q].w);
        pr[q] = eos.P(e[q].x) + (BULK ? ldo(bk, J[q]) : 0.0f);
    }
    const float v_mh = 0.5f * (comp(e[1], A + 1) + comp(e[2], A + 1));
    const float v_ph = 0.5f * (comp(e[3], A + 1) + comp(e[2], A + 1));
    const float lam_mh = max_prop_speed(p, eos, half4(e[1], e[2]), v_mh);
    const float lam_ph = max_prop_speed(p, eos, half4(e[3], e[2]), v_ph);
    const float rdx = A == 0 ? p.r_dx : (A == 1 ? p.r_dy : st.r_taudz);
    // velocity gradient (kernel_IS.cl:134-137, 301-304)
    {
        float um1[4] = {u0[1] * 1.0f, u0[1] * e[1].y, u0[1] * e[1].z, u0[1] * e[1].w};
        float ui[4] = {u0[2] * 1.0f, u0[2] * e[2].y, u0[2] * e[2].z, u0[2] * e[2].w};
        float up1[4] = {u0[3] * 1.0f, u0[3] * e[3].y, u0[3] * e[3].z, u0[3] * e[3].w};
#pragma unroll
        for (int c = 0; c < 4; c++) {
            float d = slope(um1[c], ui[c], up1[c]);
            if (A == 0) ud[c] = d * p.r_dx;
            else if (A == 1) ud[c] = d * p.r_dy;
            else {
                float chr = ((c == 0 ? e[2].w : (c == 3 ? 1.0f : 0.0f)) * u0[2]) * st.r_tau;
                ud[c] = d * st.r_taudz + chr;
            }
        }
    }
    // Israel-Stewart flux of u^tau pi^{mu nu} (and u^tau Pi)
#pragma unroll
    for (int mn = 0; mn < 10; mn++) {
        const float *pl = pi + (int64_t)mn * p.n;
        float k = kt1d_real(u0[0] * ldo(pl, J[0]), u0[1] * ldo(pl, J[1]), u0[2] * ldo(pl, J[2]),
                            u0[3] * ldo(pl, J[3]), u0[4] * ldo(pl, J[4]), v_mh, v_ph, lam_mh, lam_ph);
        s_is[mn] = s_is[mn] - k * rdx;
    }
    if (BULK) {
        float k = kt1d_real(u0[0] * ldo(bk, J[0]), u0[1] * ldo(bk, J[1]), u0[2] * ldo(bk, J[2]),
                            u0[3] * ldo(bk, J[3]), u0[4] * ldo(bk, J[4]), v_mh, v_ph, lam_mh, lam_ph);
        bs = bs - k * rdx;
    }
    // KT flux of tau (T0^{tau mu} + pi^{tau mu}) (kernel_visc.cl:60-127)
    const float tau = st.tau;
    float src[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float Q[5], p0[5], pa[5];
#pragma unroll
        for (int i = 0; i < 5; i++) {
            float4 T = t0m_u(e[i], pr[i], u0[i]);
            p0[i] = ldo(pi + (int64_t)PI_ROW[0][q] * p.n, J[i]);
            pa[i] = ldo(pi + (int64_t)PI_ROW[A + 1][q] * p.n, J[i]);
            Q[i] = tau * (comp(T, q) + p0[i]);
        }
        const float g0 = q == 0 ? 1.0f : 0.0f;
        const float ga = q == A + 1 ? -1.0f : 0.0f;
        float DA0 = slope(Q[1], Q[2], Q[3]);
        float DA1 = slope(Q[2], Q[3], Q[4]);
        float AL = Q[2] + 0.5f * DA0;
        float AR = Q[3] - 0.5f * DA1;
        float pr_half = 0.5f * (pr[2] + pr[3]);
        float p0h = 0.5f * (p0[2] + p0[3]);
        float pih = 0.5f * (pa[2] + pa[3]);
        float Jp = (AR + pr_half * tau * g0 - tau * p0h) * v_ph - pr_half * tau * ga + tau * pih;
        float Jm = (AL + pr_half * tau * g0 - tau * p0h) * v_ph - pr_half * tau * ga + tau * pih;
        float sq = 0.5f * (Jp + Jm) - 0.5f * lam_ph * (AR - AL);
        DA1 = DA0;
        DA0 = slope(Q[0], Q[1], Q[2]);
        AL = Q[1] + 0.5f * DA0;
        AR = Q[2] - 0.5f * DA1;
        pr_half = 0.5f * (pr[1] + pr[2]);
        p0h = 0.5f * (p0[1] + p0[2]);
        pih = 0.5f * (pa[1] + pa[2]);
        Jp = (AR + pr_half * tau * g0 - tau * p0h) * v_mh - pr_half * tau * ga + tau * pih;
        Jm = (AL + pr_half * tau * g0 - tau * p0h) * v_mh - pr_half * tau * ga + tau * pih;
        sq -= 0.5f * (Jp + Jm) - 0.5f * lam_mh * (AR - AL);
        src[q] = sq;
    }
#pragma unroll
    for (int m = 0; m < 4; m++) s_kt[m] = s_kt[m] - src[m] * rdx;
}

// STEP 1: ev_st = ev1 = d_ev[1] (ev_n), pi_st = pi1 = pi_n, ev2 = stale d_ev[2];
//         writes pi*, Pi*, ev*, IS_src and Src (stage-1 sources), goodcell.
// STEP 2: ev_st = ev2 = ev*, pi_st = pi*, ev1 = ev_n, pi1 = pi_n;
//         writes pi_{n+1}, Pi_{n+1}, ev_{n+1}, udiff, goodcell (the step path does not
//         write the stage-2 sums of the sources back).
#ifndef CV_STAGE_WAVES
#define CV_STAGE_WAVES 1
#endif
template <bool BULK, int STEP>
__global__ __launch_bounds__(256, CV_STAGE_WAVES) void k_visc_stage(PhysP p, EosP E, StageP st,
        const float *__restrict__ ev_st, const float *__restrict__ pi_st, const float *__restrict__ bk_st,
        const float *ev1, const float *ev2, const float *pi1, const float *bk1,
        const float *__restrict__ udiff_in, float *IS_src, float *bsrc, float *Src,
        float *pi_out, float *bk_out, float *ev_out, float *good, float *udiff_out, JetP jet) {
    int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= p.n) return;
    Eos eos(E, E.tab);
    Cell cl = cell_of(p, I);
    const int64_t n = p.n;
    float s[10], bs = 0.0f, t[4];
    if (STEP == 1) {
#pragma unroll
        for (int m = 0; m < 10; m++) s[m] = 0.0f;
        t[0] = t[1] = t[2] = t[3] = 0.0f;
    } else {
#pragma unroll
        for (int m = 0; m < 10; m++) s[m] = IS_src[(int64_t)m * n + I];
        if (BULK) bs = bsrc[I];
        float4 q = ld4(Src, n, I);
        t[0] = q.x; t[1] = q.y; t[2] = q.z; t[3] = q.w;
    }
    const float4 ec = ld4(ev_st, n, I);
    float q[10];
#pragma unroll
    for (int m = 0; m < 10; m++) q[m] = pi_st[(int64_t)m * n + I];
    // visc_src_christoffel (kernel_IS.cl:41-67)
    {
        float uz = ec.w * gamma3(ec.y, ec.z, ec.w);
        s[0] -= (2.0f * uz * q[3]) * st.r_tau;
        s[1] -= (uz * q[6]) * st.r_tau;
        s[2] -= (uz * q[8]) * st.r_tau;
        s[3] -= (uz * (q[0] + q[9])) * st.r_tau;
        s[6] -= (uz * q[1]) * st.r_tau;
        s[8] -= (uz * q[2]) * st.r_tau;
        s[9] -= (2.0f * uz * q[3]) * st.r_tau;
    }
    // kt_src_christoffel (kernel_visc.cl:11-53)
    {
        float u0 = gamma3(ec.y, ec.z, ec.w);
        float bulkpi = BULK ? bk_st[I] : 0.0f;
        float pressure = eos.P(ec.x) + bulkpi;
        float Tzz = (ec.x + pressure) * u0 * u0 * ec.w * ec.w + pressure + q[9];
        float Ttz = (ec.x + pressure) * u0 * u0 * ec.w + q[3];
        t[0] = t[0] - Tzz; t[1] = t[1] - 0.0f; t[2] = t[2] - 0.0f; t[3] = t[3] - Ttz;
        if (p.flags & 1) {
            float Ttt = (ec.x + pressure) * u0 * u0 - pressure;
            float Ttx = (ec.x + pressure) * u0 * u0 * ec.y;
            float Tty = (ec.x + pressure) * u0 * u0 * ec.z;
            t[0] = t[0] + (Tzz + Ttt); t[1] = t[1] + Ttx; t[2] = t[2] + Tty; t[3] = t[3] + 3.0f * Ttz;
        }
    }
    float udx[4], udy[4], udz[4];
    // scheduling fences keep each axis' stencil loads within its own phase (register pressure)
    CV_PHASE_FENCE();
    visc_axis<0, BULK>(p, eos, st, cl, ev_st, pi_st, bk_st, s, bs, t, udx);
    CV_PHASE_FENCE();
    visc_axis<1, BULK>(p, eos, st, cl, ev_st, pi_st, bk_st, s, bs, t, udy);
    CV_PHASE_FENCE();
    visc_axis<2, BULK>(p, eos, st, cl, ev_st, pi_st, bk_st, s, bs, t, udz);
    CV_PHASE_FENCE();
    if (jet.on & st.step) jet_add(t, jet_deposit(p, eos, jet, st.tau, ev_st, cl));

    // ---- update_pimn (kernel_IS.cl:501-709) ----
    const float4 e1 = STEP == 1 ? ec : ld4(ev1, n, I);
    const float4 e2 = STEP == 2 ? ec : ld4(ev2, n, I);
    float g1 = gamma3(e1.y, e1.z, e1.w);
    float u_old[4] = {g1 * 1.0f, g1 * e1.y, g1 * e1.z, g1 * e1.w};
    float u_new[4], udt[4];
    if (STEP == 1) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            float ud_ = udiff_in[(int64_t)c * n + I];
            udt[c] = ud_ * p.r_dt;
            u_new[c] = u_old[c] + ud_;
        }
    } else {
        float g2 = gamma3(e2.y, e2.z, e2.w);
        u_new[0] = g2 * 1.0f; u_new[1] = g2 * e2.y; u_new[2] = g2 * e2.z; u_new[3] = g2 * e2.w;
#pragma unroll
        for (int c = 0; c < 4; c++) udt[c] = (u_new[c] - u_old[c]) * p.r_dt;
    }
    float da[4][4];
#pragma unroll
    for (int c = 0; c < 4; c++) { da[c][0] = udt[c]; da[c][1] = udx[c]; da[c][2] = udy[c]; da[c][3] = udz[c]; }
    float u[4], DU[4];
#pragma unroll
    for (int c = 0; c < 4; c++) u[c] = STEP == 1 ? u_old[c] : u_new[c];
#pragma unroll
    for (int c = 0; c < 4; c++) DU[c] = dot4(u, da[c]);
    const float ed_step = STEP == 1 ? e1.x : e2.x;
    const float theta = udt[0] + udx[1] + udy[2] + udz[3];
    const float4 cpTs = eos.raw(ed_step);
    const float temperature = cpTs.z;
    const float local_etaos = etaos(p, temperature);
    float etav = local_etaos * cpTs.w * CV_HBARC;
    float one_over_taupi = temperature * rcp(3.0f * rmax(CV_ACU, local_etaos) * CV_HBARC);
    float coef_pipi = 0.0f;
    const bool gubser = p.flags & 2;
    if (gubser) {
        float q4 = sqrtf(sqrtf(ed_step));
        float ETAOS = p.etaos_ymin;
        float etavH = (ETAOS * 4.0f) * p.r_3;
        float taupiH = (p.lam1 * p.lam1 * etavH) * p.r_3;
        one_over_taupi = rcp(taupiH * rcp(q4));
        coef_pipi = p.lam1 * rcp(ed_step);
        etav = ((ETAOS * 4.0f) * p.r_3) * (sqrtf(ed_step) * q4);
    }
    const bool cpl = (p.flags & (CV_FLAG_PIMN_OMEGA | CV_FLAG_OMEGA_OMEGA)) != 0;
    float Om[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (cpl) omega_cell(udt, udx, udy, udz, u, Om);
    const float gd[4] = {1.0f, -1.0f, -1.0f, -1.0f};
    float max_abs = 0.0f;
    float out[10];
    const float rden = rcp(u_new[0] + p.dt * one_over_taupi);
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
#pragma unroll
        for (int nu = mu; nu < 4; nu++) {
            float gmunu = mu == nu ? gd[mu] : 0.0f;
            float sigma = gd[mu] * da[nu][mu] + gd[nu] * da[mu][nu] - (u[mu] * DU[nu] + u[nu] * DU[mu])
                          - 2.0f / 3.0f * (gmunu - u[mu] * u[nu]) * theta;
            int mn = IDX(mu, nu);
            if (u[0] > 10000.0f) sigma = 0.0f;
            float piNS = etav * sigma;
            float pi_old = (STEP == 1 ? q[mn] : pi1[(int64_t)mn * n + I]) * u_old[0];
            float src = (-q[mn] * theta) * p.r_3 - (q[mn] * u[0]) * st.r_tau;
#pragma unroll
            for (int l = 0; l < 4; l++)
                src -= (u[mu] * q[IDX(nu, l)] + u[nu] * q[IDX(mu, l)]) * DU[l] * gd[l];
            if (gubser)
                src -= one_over_taupi * coef_pipi * (pipi(0, mu, nu, q, u) - pipi(1, mu, nu, q, u)
                                                     - pipi(2, mu, nu, q, u) - pipi(3, mu, nu, q, u));
            if (cpl) src = couple_src(src, p.flags, mu, nu, q, Om, u, one_over_taupi);
            s[mn] += src;
            pi_old = (pi_old + (s[mn] * p.dt) * st.r_step + p.dt * one_over_taupi * piNS) * rden;
            if (fabsf(pi_old) > max_abs) max_abs = fabsf(pi_old);
            out[mn] = pi_old;
        }
    }
    float bout = 0.0f, b_start = 0.0f;
    if (BULK) {
        float zs = zetaos(p, temperature);
        float zetav = zs * cpTs.w * CV_HBARC;
        float oot_b = temperature * rcp(3.0f * rmax(CV_ACU, zs) * CV_HBARC);
        float b2 = bk_st[I];
        float bsrc_ = (b2 * theta) * p.r_3 - (b2 * u[0]) * st.r_tau;
        bs += bsrc_;
        b_start = STEP == 1 ? b2 : bk1[I];
        float b_old = b_start * u_old[0];
        bout = (b_old + (bs * p.dt) * st.r_step - p.dt * oot_b * zetav * theta) * rcp(u_new[0] + p.dt * oot_b);
        bout = bulk_regulate(bout, eos.P(e2.x), e2.x);
        if (fabsf(bout) > max_abs) max_abs = fabsf(bout);
    }
    {
        float pre = eos.P(e2.x);
        enthalpy_cap<BULK>(p.zmax, out, bout, max_abs, e2.x, pre);
        float T00 = (e2.x + pre) * u_new[0] * u_new[0] - pre;
        float gc = 1.0f;
        if (max_abs > T00) {
#pragma unroll
            for (int m = 0; m < 10; m++) out[m] = 0.0f;
            bout = 0.0f;
            gc = 0.0f;
        }
        good[I] = gc;
    }
#pragma unroll
    for (int m = 0; m < 10; m++) pi_out[(int64_t)m * n + I] = out[m];
    if (BULK) bk_out[I] = bout;
    if (STEP == 1) {
#pragma unroll
        for (int m = 0; m < 10; m++) IS_src[(int64_t)m * n + I] = s[m];
        if (BULK) bsrc[I] = bs;
        st4(Src, n, I, make_float4(t[0], t[1], t[2], t[3]));
    }

    // ---- update_ev (kernel_visc.cl:322-393): start point ev_n, pi_n; new pi in registers ----
    float pressure = eos.P(e1.x) + (BULK ? b_start : 0.0f);
    float T0m[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        float g0 = c == 0 ? 1.0f : 0.0f;
        float pim0 = STEP == 1 ? q[c] : pi1[(int64_t)c * n + I];
        T0m[c] = ((e1.x + pressure) * g1 * u_old[c] - pressure * g0 + pim0) * st.old_time;
    }
#pragma unroll
    for (int c = 0; c < 4; c++) T0m[c] = T0m[c] + (t[c] * p.dt) * st.r_step - out[c] * st.tau;
    float T00 = rmax(CV_ACU, T0m[0]) * st.r_tau;
    float T01 = fabsf(T0m[1]) < CV_ACU ? 0.0f : T0m[1] * st.r_tau;
    float T02 = fabsf(T0m[2]) < CV_ACU ? 0.0f : T0m[2] * st.r_tau;
    float T03 = fabsf(T0m[3]) < CV_ACU ? 0.0f : T0m[3] * st.r_tau;
    float M = sqrtf(T01 * T01 + T02 * T02 + T03 * T03);
    if (M > T00) {
        float sc = (0.999f * T00) * rcp(M);
        T01 *= sc; T02 *= sc; T03 *= sc;
        M = 0.999f * T00;
    }
    float ed_find;
    if ((p.flags & 8) && M > p.fp64_vmin * T00)
        ed_find = (float)root_newton_d(E, E.tab, (double)T00, (double)M, (double)bout, BULK);
    else
        ed_find = root_newton(eos, T00, M, bout, BULK);
    if (ed_find < CV_ACU) { T01 = 0.0f; T02 = 0.0f; T03 = 0.0f; }
    ed_find = rmax(0.0f, ed_find);
    float pr = eos.P(ed_find) + bout;
    if (BULK) pr = rmax(pr, 0.0f);
    float repv = rcp(rmax(CV_ACU, T00 + pr));
    float4 en = make_float4(ed_find, T01 * repv, T02 * repv, T03 * repv);
    st4(ev_out, n, I, en);
    if (STEP == 2) {
        float gn = gamma3(en.y, en.z, en.w);
        float un[4] = {gn * 1.0f, gn * en.y, gn * en.z, gn * en.w};
#pragma unroll
        for (int c = 0; c < 4; c++) udiff_out[(int64_t)c * n + I] = un[c] - u_old[c];
    }
}

// visc_initialize (kernel_IS.cl:10-36)
__global__ void k_visc_init(PhysP p, EosP E, float tau, float r_tau, const float *__restrict__ ev,
                            float *pi1, float *good, float *udiff) {
    int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= p.n) return;
    Eos eos(E, E.tab);
    float4 cpTs = eos.raw(ev[I]);
    float etav = etaos(p, cpTs.z) * cpTs.w * CV_HBARC;
    float tmp = (2.0f / 3.0f * etav) * r_tau;
    pi1[4 * p.n + I] = tmp;
    pi1[7 * p.n + I] = tmp;
    pi1[9 * p.n + I] = -2.0f * tmp;
    for (int c = 0; c < 4; c++) udiff[c * p.n + I] = 0.0f;
    good[I] = 1.0f;
}

// get_udiff (kernel_IS.cl:714-724)
__global__ void k_udiff(PhysP p, const float *__restrict__ ev0, const float *__restrict__ ev1, float *udiff) {
    int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= p.n) return;
    float4 a = ld4(ev0, p.n, I), b = ld4(ev1, p.n, I);
    float ga = gamma3(a.y, a.z, a.w), gb = gamma3(b.y, b.z, b.w);
    float ua[4] = {ga * 1.0f, ga * a.y, ga * a.z, ga * a.w};
    float ub[4] = {gb * 1.0f, gb * b.y, gb * b.z, gb * b.w};
    for (int c = 0; c < 4; c++) udiff[c * p.n + I] = ub[c] - ua[c];
}

// pizz_o_ep (kernel_checkpi.cl:5-27, CLVisc.check_pizz): for eta slice K = blockIdx.x the
// mean of pi^{eta eta}/(e + P) over the cells with e + P > 0.133.  The 64 lanes evaluate 64
// consecutive cells of the slice (cell c = i NY + j, so c runs over i, then j) and lane 0 adds
// their terms in that order, the reference's one-thread sum bit for bit.
__global__ __launch_bounds__(64) void k_pizz_o_ep(PhysP p, EosP E, const float *__restrict__ ev,
                                                  const float *__restrict__ pi, float *out) {
    Eos eos(E, E.tab);
    const int64_t n = p.n, nij = (int64_t)p.nx * p.ny;
    const int64_t K = blockIdx.x;
    __shared__ float s_q[64];
    __shared__ int s_ok[64];
    float ratio = 0.0f;
    int num_of_cells = 0;
    for (int64_t b = 0; b < nij; b += 64) {
        const int64_t c = b + threadIdx.x;
        float q = 0.0f;
        int ok = 0;
        if (c < nij) {
            const int64_t I = c * p.nz + K;
            const float e = ev[I];
            const float ep = e + eos.P(e);
            if (ep > 0.133f) {
                q = pi[9 * n + I] * rcp(ep);
                ok = 1;
            }
        }
        s_q[threadIdx.x] = q;
        s_ok[threadIdx.x] = ok;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int64_t m = nij - b < 64 ? nij - b : 64;
            for (int64_t k = 0; k < m; k++)
                if (s_ok[k]) {
                    ratio += s_q[k];
                    num_of_cells++;
                }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[K] = ratio * rcp((float)num_of_cells);
}

// max of e (reduction_stage1 + host max, kernel_reduction.cl:7-39): block max by
// wavefront shuffles, one atomic per block on the order-preserving integer image.
__device__ __forceinline__ unsigned int f2ord(float f) {
    unsigned int u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__global__ __launch_bounds__(256) void k_max_ed(int64_t n, const float *__restrict__ e, unsigned int *out) {
    float m = -3.402823466e+38f;
    for (int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; I < n; I += (int64_t)gridDim.x * blockDim.x)
        m = e[I] > m ? e[I] : m;
    for (int off = 32; off > 0; off >>= 1) {
        float o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    __shared__ float wm[4];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wm[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) m = wm[q] > m ? wm[q] : m;
        atomicMax(out, f2ord(m));
    }
}

// FP64 energy balance of the conserved form Q^tau = tau (T^tautau + pi^tautau), whose only
// source besides the flux divergence is -T~^etaeta (kt_src_christoffel, kernel_visc.cl:11-40):
// part[b] = (sum tau T^tautau, sum T~^etaeta) over the cells of block b, in double from the
// FP32 fields (P from the FP32 table; bulk Pi added to P).  Fixed grid and order: the sums
// are reproducible run to run.  k_energy_final adds the partials in a fixed order.
#define CV_ESUM_BLOCKS 1024
__global__ __launch_bounds__(256) void k_energy_sum(PhysP p, EosP E, const float *__restrict__ ev,
                                                    const float *__restrict__ pi, const float *__restrict__ bk,
                                                    double tau, double *part) {
    Eos eos(E, E.tab);
    const int64_t n = p.n;
    double sE = 0.0, sW = 0.0;
    for (int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; I < n; I += (int64_t)gridDim.x * blockDim.x) {
        const float e = ev[I];
        const double vx = ev[n + I], vy = ev[2 * n + I], vz = ev[3 * n + I];
        const double pr = (double)eos.P(e) + (bk ? (double)bk[I] : 0.0);
        const double g2 = 1.0 / (1.0 - (vx * vx + vy * vy + vz * vz));
        const double w = ((double)e + pr) * g2;
        const double p00 = pi ? (double)pi[I] : 0.0, p33 = pi ? (double)pi[9 * n + I] : 0.0;
        sE += tau * (w - pr + p00);
        sW += w * vz * vz + pr + p33;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sE += __shfl_xor(sE, off, 64);
        sW += __shfl_xor(sW, off, 64);
    }
    __shared__ double s[2][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { s[0][wv] = sE; s[1][wv] = sW; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = ((s[0][0] + s[0][1]) + s[0][2]) + s[0][3];
        part[2 * blockIdx.x + 1] = ((s[1][0] + s[1][1]) + s[1][2]) + s[1][3];
    }
}

// one workgroup of 256: each thread adds partials q = tid, tid + 256, ... in order, then a
// fixed tree over the 256 (the same order on every call)
__global__ __launch_bounds__(256) void k_energy_final(const double *__restrict__ part, int nb, double dV,
                                                      double *out) {
    __shared__ double sa[256], sb[256];
    double a = 0.0, b = 0.0;
    for (int q = threadIdx.x; q < nb; q += 256) { a += part[2 * q]; b += part[2 * q + 1]; }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sa[threadIdx.x] = sa[threadIdx.x] + sa[threadIdx.x + w];
            sb[threadIdx.x] = sb[threadIdx.x] + sb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sa[0] * dV;
        out[1] = sb[0] * dV;
    }
}

// BulkInfo observables per eta slice (bulkinfo.py:203-260, ecc_vs_rapidity :247-260): for
// every k, sums over the (x, y) plane of Txx - Tyy, Txx + Tyy and T^tx (eccp), u0 s (total
// entropy), u0 e |v_T| and u0 e (<v_r>), with e floored at 1e-10 and v^2 > 1 replaced by
// 0.999999 as the reference's NumPy code does; P, s from the FP32 table, sums in double.
// Block b covers rows (i, j) [b R, (b + 1) R); thread k of the block walks them at its k, so
// a warp reads consecutive cells.  part[b][k][q]; k_slice_final adds them in a fixed order.
#define CV_SLICE_Q 10
__global__ __launch_bounds__(128) void k_slice_sums(PhysP p, EosP E, const float *__restrict__ ev,
                                                    int64_t rows_per_block, double *part) {
    Eos eos(E, E.tab);
    const int64_t n = p.n, nrow = (int64_t)p.nx * p.ny;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
    const int64_t r1 = r0 + rows_per_block < nrow ? r0 + rows_per_block : nrow;
    for (int k = threadIdx.x; k < p.nz; k += blockDim.x) {
        double s[CV_SLICE_Q] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int64_t r = r0; r < r1; r++) {
            const int64_t I = r * p.nz + k;
            const float ef = ev[I] < 1.0e-10f ? 1.0e-10f : ev[I];
            const double e = ef, vx = ev[n + I], vy = ev[2 * n + I], vz = ev[3 * n + I];
            const double pr = (double)eos.P(ef);
            double vr2 = vx * vx + vy * vy + vz * vz;
            if (vr2 > 1.0) vr2 = 0.999999;
            const double u0 = 1.0 / sqrt(1.0 - vr2);
            const double w = (e + pr) * u0 * u0;
            const double txx = w * vx * vx + pr, tyy = w * vy * vy + pr;
            s[0] += txx - tyy;
            s[1] += txx + tyy;
            s[2] += w * vx;
            s[3] += u0 * (double)eos.S(ef);
            s[4] += u0 * e * sqrt(vx * vx + vy * vy);
            s[5] += u0 * e;
            // total_energy_and_entropy (kernel_bulkinfo.cl:18-46): unfloored e, gamma with
            // 1 - v^2 floored at 1e-7 (helper.h:46-48); the lab-frame energy of the slice is
            // cosh(eta) (s[6] - s[8]) + sinh(eta) s[7]
            const float er = ev[I];
            const double prr = (double)eos.P(er);
            double g2 = 1.0 - (vx * vx + vy * vy + vz * vz);
            g2 = g2 > 1.0e-7 ? g2 : 1.0e-7;
            const double ug = 1.0 / sqrt(g2);
            const double wr = ((double)er + prr) * ug * ug;
            s[6] += wr;
            s[7] += wr * vz;
            s[8] += prr;
            s[9] += ug * (double)eos.S(er);
        }
#pragma unroll
        for (int q = 0; q < CV_SLICE_Q; q++) part[((int64_t)blockIdx.x * p.nz + k) * CV_SLICE_Q + q] = s[q];
    }
}

// one workgroup of 256 per output: each thread sums partials b = tid, tid + 256, ... in
// order, then a fixed tree over the 256 (the same order on every call)
__global__ __launch_bounds__(256) void k_slice_final(const double *__restrict__ part, int nb, int nz, double *out) {
    __shared__ double s[256];
    const int t = blockIdx.x;                       // output index, < nz * CV_SLICE_Q
    double a = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) a += part[(int64_t)b * nz * CV_SLICE_Q + t];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[t] = s[0];
}

// out[q][c] = field plane c at cell idx[q] (SoA planes -> AoS rows); idx checked on the host
__global__ __launch_bounds__(256) void k_gather(int64_t n, int ncomp, const float *__restrict__ f,
                                                const int64_t *__restrict__ idx, int64_t m, float *out) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    int64_t I = idx[q];
    for (int c = 0; c < ncomp; c++) out[q * ncomp + c] = f[(int64_t)c * n + I];
}

}  // namespace cv
