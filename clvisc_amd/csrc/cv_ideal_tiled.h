// cv_ideal_tiled.h -- x-marching tiled stage kernel of the ideal step (gfx950).
//
// CLIdeal.stepUpdate (pyvisc/ideal.py:211-255): kt_src_christoffel, kt_src_along{x,y,z}
// and update_ev (kernel_ideal.cl:3-231) of one RK stage in one launch, with the work
// decomposition of the viscous marching kernel (cv_tiled.h): a workgroup owns a ty x tz
// tile of the (y, eta) plane and walks a chunk of x-planes.  Per plane the per-cell
// quantities (e, v, P, T^{tau mu}) of the tile and its 2-cell y/z halo go to LDS once,
// every y and z face is evaluated once, and the x face F(i+1/2) comes from the LDS cell i
// and cells i+1, i+2 of the own column, with F(i-1/2), the slopes of cell i and the
// quantities of cell i+1 carried from the previous plane.  A face value is the expression
// kt1d evaluates as its first half at cell i and as its second half at cell i+1
// (helper.h:179-242): the same function of cells i-1 .. i+2 (the half state, the pressure
// average and the face velocity are sums of the two cells, which commute), so evaluating
// it once gives the reference's value bit for bit.
#pragma once
#include "cv_tiled.h"

namespace cv {

// per-cell fields of the ideal tile in LDS
enum { CI_E = 0, CI_VX, CI_VY, CI_VZ, CI_P, CI_T0, CI_N = CI_T0 + 4 };

#define CV_ITILE_THREADS 256

template <class EOS>
__device__ __forceinline__ void ideal_prims(const EOS &eos, float4 e, float out[CI_N]) {
    const float pr = eos.P(e.x);
    const float4 T = t0m(e, pr);                     // kt1d's t0m(c, P(c.x))
    out[CI_E] = e.x; out[CI_VX] = e.y; out[CI_VY] = e.z; out[CI_VZ] = e.w;
    out[CI_P] = pr;
    out[CI_T0] = T.x; out[CI_T0 + 1] = T.y; out[CI_T0 + 2] = T.z; out[CI_T0 + 3] = T.w;
}

// KT flux of T^{tau mu} at the face between cells L and R along axis A, from the cells
// Lm = L-1, L, R, Rp = R+1 (kt1d, helper.h:179-242).  HAVE_DAL: the slopes of cell L are
// given in dal (Lm is not read); WANT_DAR: the slopes of cell R are returned in dar.
template <int A, bool HAVE_DAL = false, bool WANT_DAR = false, class EOS, typename G>
__device__ __forceinline__ void ideal_face(const PhysP &p, const EOS &eos, const G &g, int Lm, int L, int R, int Rp,
                                           float out[4], const float *dal = nullptr, float *dar = nullptr) {
    const float4 eL = make_float4(g(CI_E, L), g(CI_VX, L), g(CI_VY, L), g(CI_VZ, L));
    const float4 eR = make_float4(g(CI_E, R), g(CI_VX, R), g(CI_VY, R), g(CI_VZ, R));
    float AL[4], AR[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float DAL = HAVE_DAL ? dal[q] : slope(g(CI_T0 + q, Lm), g(CI_T0 + q, L), g(CI_T0 + q, R));
        const float DAR = slope(g(CI_T0 + q, L), g(CI_T0 + q, R), g(CI_T0 + q, Rp));
        if (WANT_DAR) dar[q] = DAR;
        AL[q] = g(CI_T0 + q, L) + 0.5f * DAL;
        AR[q] = g(CI_T0 + q, R) - 0.5f * DAR;
    }
    const float pr_half = 0.5f * (g(CI_P, R) + g(CI_P, L));
    const float vh = 0.5f * (comp(eL, A + 1) + comp(eR, A + 1));
    const float lam = max_prop_speed(p, eos, half4(eL, eR), vh);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float g0 = q == 0 ? 1.0f : 0.0f;
        const float ga = q == A + 1 ? -1.0f : 0.0f;
        const float Jp = (AR[q] + pr_half * g0) * vh - pr_half * ga;
        const float Jm = (AL[q] + pr_half * g0) * vh - pr_half * ga;
        out[q] = 0.5f * (Jp + Jm) - 0.5f * lam * (AR[q] - AL[q]);
    }
}

// One RK stage of the ideal step (k_ideal_stage without the test flags, the FP64 root,
// the learned seed and the iteration counts, which take the per-cell kernel).
template <int EK>
__global__ __launch_bounds__(CV_ITILE_THREADS) void k_ideal_tiled(PhysP p, EosP E, StageP st, TileP tp,
                                                                  const float *__restrict__ ev_st,
                                                                  const float *__restrict__ ev1, float *Src,
                                                                  float *ev_out) {
    __shared__ float s_c[CI_N][CV_TILE_MAXC];
    __shared__ float s_f[4][CV_TILE_MAXF];
    // octave table: its pressure column in LDS, as in the viscous marching kernel (the host
    // launches this kernel only for tables of at most CV_PTAB_LDS nodes)
    __shared__ float s_pt[EK == 1 ? CV_PTAB_LDS : 1];
    const auto eos = [&]() {
        if constexpr (EK == 1) return EosOctP(E, E.tab, s_pt);
        else return EosT<EK>(E, E.tab);
    }();
    const int tid = threadIdx.x;
    const uint32_t nb = p.nb;
    const float tau = st.tau;
    const int ty = tp.ty, tz = tp.tz, HY = tp.HY, HZ = tp.HZ, hy = tp.hy, hz = tp.hz;
    const int my = tid / tz, mz = tid - (tid / tz) * tz;
    const bool lane_in_tile = tid < ty * tz;
    const int64_t work = (tp.xcd > 0 && (int64_t)gridDim.x == tp.nwork) ? xcd_work(blockIdx.x, tp.nwork, tp.xcd)
                                                                         : (int64_t)blockIdx.x;
    if (work >= tp.nwork) return;                      // uniform over the workgroup
    if constexpr (EK == 1) {
        for (int n = tid; n <= E.nlast; n += CV_ITILE_THREADS) s_pt[n] = E.tab[n].y;
        __syncthreads();
    }
    const int kt = (int)(work % tp.ntk);
    const int64_t r = work / tp.ntk;
    const int jt = (int)(r % tp.ntj);
    int ibeg, iend;
    march_planes(tp, p.nx, (int)(r / tp.ntj), ibeg, iend);
    if (ibeg >= iend) return;                          // no planes (uniform over the workgroup)
    const int j0 = jt * ty, k0 = kt * tz;
    const int j = j0 + my, k = k0 + mz;
    const bool active = lane_in_tile && j < p.ny && k < p.nz;
    const int cc = (my + hy) * HZ + (mz + hz);
    // carried along x: F(i-1/2), slopes of cell i, the quantities of cell i+1
    float xFm[4], xDAL[4], xr1[CI_N];
    if (active) {
        const uint32_t I0 = (uint32_t)((int64_t)ibeg * p.sx) + (uint32_t)j * (uint32_t)p.sy + (uint32_t)k;
        float xc[4][CI_N];                               // cells ibeg-2 .. ibeg+1
#pragma unroll
        for (int w = 0; w < 4; w++) {
            int ii = ibeg + w - 2;
            ii = ii < 0 ? 0 : (ii > p.nx - 1 ? p.nx - 1 : ii);
            ideal_prims(eos, ld4p(ev_st, (I0 + (uint32_t)(ii - ibeg) * (uint32_t)p.sx) * 4u, nb), xc[w]);
        }
        auto gx = [&](int f, int c) { return xc[c][f]; };
        ideal_face<0, false, true>(p, eos, gx, 0, 1, 2, 3, xFm, nullptr, xDAL);
#pragma unroll
        for (int f = 0; f < CI_N; f++) xr1[f] = xc[3][f];
    }
    for (int i = ibeg; i < iend; i++) {
        const uint32_t plane = (uint32_t)((int64_t)i * p.sx);
        // ---- per-cell quantities of tile + y/z halo into LDS ----
        for (int c = tid; c < HY * HZ; c += CV_ITILE_THREADS) {
            const int ly = c / HZ, lz = c - (c / HZ) * HZ;
            const bool hy_out = ly < hy || ly >= hy + ty, hz_out = lz < hz || lz >= hz + tz;
            if (hy_out && hz_out) continue;
            int jj = j0 + ly - hy, kk = k0 + lz - hz;
            jj = jj < 0 ? 0 : (jj > p.ny - 1 ? p.ny - 1 : jj);
            kk = kk < 0 ? 0 : (kk > p.nz - 1 ? p.nz - 1 : kk);
            float out[CI_N];
            ideal_prims(eos, ld4p(ev_st, (plane + (uint32_t)jj * (uint32_t)p.sy + (uint32_t)kk) * 4u, nb), out);
#pragma unroll
            for (int f = 0; f < CI_N; f++) s_c[f][c] = out[f];
        }
        __syncthreads();
        const uint32_t I = plane + (uint32_t)j * (uint32_t)p.sy + (uint32_t)k;
        const uint32_t bo = I * 4u;
        auto gl = [&](int f, int c) { return s_c[f][c]; };
        float s[4];
        if (active) {
            // kt_src_christoffel (kernel_ideal.cl:3-36)
            if (st.step == 1) {
                s[0] = s[1] = s[2] = s[3] = 0.0f;
            } else {
                const float4 t = ld4p(Src, bo, nb);
                s[0] = t.x; s[1] = t.y; s[2] = t.z; s[3] = t.w;
            }
            {
                const float4 e = make_float4(s_c[CI_E][cc], s_c[CI_VX][cc], s_c[CI_VY][cc], s_c[CI_VZ][cc]);
                const float pressure = s_c[CI_P][cc];
                const float u0 = gamma3(e.y, e.z, e.w);
                const float Tzz = (e.x + pressure) * u0 * u0 * e.w * e.w + pressure;
                const float Ttz = (e.x + pressure) * u0 * u0 * e.w;
                s[0] = s[0] - Tzz; s[1] = s[1] - 0.0f; s[2] = s[2] - 0.0f; s[3] = s[3] - Ttz;
            }
            // ---- x: F(i+1/2) from cell i (LDS), i+1 (carried), i+2 ----
            float xr2[CI_N];
            {
                int ii = i + 2;
                ii = ii > p.nx - 1 ? p.nx - 1 : ii;
                ideal_prims(eos, ld4p(ev_st, (I + (uint32_t)(ii - i) * (uint32_t)p.sx) * 4u, nb), xr2);
            }
            auto gx = [&](int f, int c) { return c == 1 ? s_c[f][cc] : (c == 2 ? xr1[f] : xr2[f]); };
            float Fp[4], dar[4];
            ideal_face<0, true, true>(p, eos, gx, 0, 1, 2, 3, Fp, xDAL, dar);
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const float kk = Fp[m] - xFm[m];
                s[m] = s[m] - (st.tau * kk) * p.r_dx;
                xFm[m] = Fp[m];
                xDAL[m] = dar[m];
            }
#pragma unroll
            for (int f = 0; f < CI_N; f++) xr1[f] = xr2[f];
        }
        // ---- y faces: face f in [0, ty] lies between tile rows f-1 and f ----
        if (hy > 0) {
            for (int fc = tid; fc < (ty + 1) * tz; fc += CV_ITILE_THREADS) {
                const int f = fc / tz, z = fc - (fc / tz) * tz;
                const int Lc = (f - 1 + hy) * HZ + (z + hz);
                float F[4];
                ideal_face<1>(p, eos, gl, Lc - HZ, Lc, Lc + HZ, Lc + 2 * HZ, F);
#pragma unroll
                for (int m = 0; m < 4; m++) s_f[m][fc] = F[m];
            }
        }
        __syncthreads();
        if (active) {
            if (hy > 0) {
                const int fp_ = (my + 1) * tz + mz, fm_ = my * tz + mz;
#pragma unroll
                for (int m = 0; m < 4; m++) s[m] = s[m] - (st.tau * (s_f[m][fp_] - s_f[m][fm_])) * p.r_dy;
            } else {
                // ny == 1: both faces are the same expression of the same cells -> 0
#pragma unroll
                for (int m = 0; m < 4; m++) s[m] = s[m] - (st.tau * 0.0f) * p.r_dy;
            }
        }
        __syncthreads();
        // ---- z faces: face f in [0, tz] lies between tile columns f-1 and f ----
        if (hz > 0) {
            for (int fc = tid; fc < ty * (tz + 1); fc += CV_ITILE_THREADS) {
                const int y = fc / (tz + 1), f = fc - (fc / (tz + 1)) * (tz + 1);
                const int Lc = (y + hy) * HZ + (f - 1 + hz);
                float F[4];
                ideal_face<2>(p, eos, gl, Lc - 1, Lc, Lc + 1, Lc + 2, F);
#pragma unroll
                for (int m = 0; m < 4; m++) s_f[m][fc] = F[m];
            }
        }
        __syncthreads();
        if (active) {
            if (hz > 0) {
                const int fp_ = my * (tz + 1) + mz + 1, fm_ = my * (tz + 1) + mz;
#pragma unroll
                for (int m = 0; m < 4; m++) s[m] = s[m] - (st.tau * (s_f[m][fp_] - s_f[m][fm_])) * st.r_taudz;
            } else {
#pragma unroll
                for (int m = 0; m < 4; m++) s[m] = s[m] - (st.tau * 0.0f) * st.r_taudz;
            }
#pragma unroll
            for (int m = 0; m < 4; m++) stp(Src, bo, nb, m, s[m]);
            // ---- update_ev (kernel_ideal.cl:165-231) ----
            const float4 e = ld4p(ev1, bo, nb);
            const float pressure = eos.P(e.x);
            const float u0 = gamma3(e.y, e.z, e.w);
            const float umu[4] = {u0 * 1.0f, u0 * e.y, u0 * e.z, u0 * e.w};
            float T0m[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float g0 = q == 0 ? 1.0f : 0.0f;
                const float t = (e.x + pressure) * u0 * umu[q] - pressure * g0;
                T0m[q] = t * st.old_time;
            }
#pragma unroll
            for (int q = 0; q < 4; q++) T0m[q] = T0m[q] + (s[q] * p.dt) * st.r_step;
            float T00 = rmax(CV_ACU, T0m[0]) * st.r_tau;
            float T01 = fabsf(T0m[1]) < CV_ACU ? 0.0f : T0m[1] * st.r_tau;
            float T02 = fabsf(T0m[2]) < CV_ACU ? 0.0f : T0m[2] * st.r_tau;
            float T03 = fabsf(T0m[3]) < CV_ACU ? 0.0f : T0m[3] * st.r_tau;
            float M = sqrtf(T01 * T01 + T02 * T02 + T03 * T03);
            const float SC = 0.999f;
            if (M > T00) {
                const float sc = (SC * T00) * rcpn(M);      // M > T00 >= 1e-7 / tau
                T01 *= sc; T02 *= sc; T03 *= sc;
                M = SC * T00;
            }
            int it = 0;
            float ed_find = root_newton_from(eos, T00, M, 0.0f, false, 0.5f * (0.0f + 1.0f), &it);
            ed_find = rmax(0.0f, ed_find);
            const float pr = eos.P(ed_find);
            const float repv = rcpn(rmax(CV_ACU, T00 + pr));   // >= 1e-7
            stp(ev_out, bo, nb, 0, ed_find);
            stp(ev_out, bo, nb, 1, T01 * repv);
            stp(ev_out, bo, nb, 2, T02 * repv);
            stp(ev_out, bo, nb, 3, T03 * repv);
        }
        __syncthreads();   // LDS reused by the next plane
    }
}

}  // namespace cv
