// cv_tiled.h -- tiled fused stage kernel of the viscous step (gfx950).
//
// One workgroup owns a tile of TY x TZ cells of one x-plane and loops (persistent) over
// tiles.  Per tile:
//   1. per-cell quantities (e, v, u^tau, P+Pi, Q = tau (T0^{tau mu} + pi^{tau mu}), pi, Pi)
//      of the tile and its 2-cell y/z halo are computed once into LDS;
//   2. x: each thread reads its own x-stencil (cells i-2..i+2) from global memory and
//      evaluates the two x faces of its cell;
//   3. y, z: every face of the tile is evaluated once from LDS, stored in LDS and
//      differenced by the two cells it separates;
//   4. update_pimn and update_ev of the cell (as in the per-cell kernel).
// A face value F(i+1/2) is the expression the reference evaluates as the first half of
// kt1d_visc / kt1d_real at cell i and again as the second half at cell i+1
// (kernel_visc.cl:86-124, helper.h:275-302): both halves are the same function of the
// cells i-1, i, i+1, i+2, so evaluating it once gives the reference's value.
// Boundary ghosts (kernel_ideal.cl:57-62) are the clamped cell indices.
// EOS: the x-marching kernel keeps the pressure column of the octave table in LDS (the
// pressure is looked up about ten times per cell and stage, the root solve's iterations in a
// dependent chain); the plane-tiled kernel copies the whole octave table to LDS.
#pragma once
#include "cv_kernels.h"
#include "cv_mlp.h"

namespace cv {

// per-cell fields held in LDS (structure of arrays)
enum { CF_E = 0, CF_VX, CF_VY, CF_VZ, CF_U0, CF_P, CF_Q0, CF_Q1, CF_Q2, CF_Q3, CF_PI0, CF_BK = CF_PI0 + 10, CF_N };

#define CV_TILE_THREADS 256
// LDS of a marching workgroup: the cells of tile + halo (21 floats each), the faces of one
// axis (15 floats each) and the pressure column of the octave table, 53.3 KB, so that 3
// workgroups fit a CU's 160 KB.  The capacities are those of the default 10 x 23 tile of the
// benchmark grid (14 x 27 = 378 cells, 11 x 23 = 253 faces).
#define CV_TILE_MAXC 378       // cells of tile + halo held in LDS
#define CV_TILE_MAXF 256       // faces of one axis held in LDS
#define CV_EOS_LDS 1600        // octave-table entries held in LDS by the plane-tiled kernel
#define CV_PTAB_LDS 1540       // octave-table pressures held in LDS by the marching kernels
#ifndef CV_TILE_EOS_IN_LDS
#define CV_TILE_EOS_IN_LDS 1
#endif

struct TileP {
    int ty, tz;                // tile interior
    int hy, hz;                // halo widths (2, or 0 for an axis of length 1)
    int HY, HZ;                // ty + 2 hy, tz + 2 hz
    int ntj, ntk;              // tiles along j, k
    int64_t ntiles;            // nx * ntj * ntk
    int xc, nchunks;           // march: x planes per chunk, chunks along x
    int64_t nwork;             // march: ntj * ntk * nchunks
    int xcd;                   // > 0: spread work items over this many XCDs in contiguous ranges
    int slabs, cps, xt;        // march, slabs > 0: x cut into `slabs` slabs of cps chunks each,
                               // the last one the slab's last xt planes
};

// x planes [ibeg, iend) of march chunk xr.  With slabs, slab s holds the planes
// [s nx / slabs, (s + 1) nx / slabs): its last chunk is its last xt planes, and the planes
// before them are shared by its cps - 1 other chunks (lengths differing by at most one).
// One slab is the work range of one XCD, so the last items each XCD runs are short ones.
__device__ __forceinline__ void march_planes(const TileP &tp, int nx, int xr, int &ibeg, int &iend) {
    if (tp.slabs > 0) {
        const int s = xr / tp.cps, c = xr - s * tp.cps, nm = tp.cps - 1;
        const int x0 = (int)(((int64_t)s * nx) / tp.slabs);
        const int x1 = (int)(((int64_t)(s + 1) * nx) / tp.slabs);
        const int xm = x1 - tp.xt;                    // end of the main chunks
        if (c == nm) {
            ibeg = xm; iend = x1;
        } else {
            ibeg = x0 + (xm - x0) * c / nm;
            iend = x0 + (xm - x0) * (c + 1) / nm;
        }
    } else {
        ibeg = xr * tp.xc;
        iend = ibeg + tp.xc < nx ? ibeg + tp.xc : nx;
    }
}

// Work item of workgroup b when the grid has one workgroup per item: the hardware hands
// workgroups to the XCDs round-robin (b % nx), so XCD x gets the contiguous range of items
// [start(x), start(x) + count(x)) and neighbouring tiles (shared y/z halos, the march's
// x-neighbours) meet in one XCD's L2.  A bijection of [0, total).
__device__ __forceinline__ int64_t xcd_work(int64_t b, int64_t total, int nx) {
    const int64_t x = b % nx, i = b / nx;
    const int64_t per = total / nx, rem = total % nx;
    return x * per + (x < rem ? x : rem) + i;
}

// per-cell quantities of one cell from its raw fields, given its u^tau and P + Pi
__device__ __forceinline__ void cell_prims_up(float tau, float4 e, const float pi[10], float bk, float u0, float pr,
                                              float out[CF_N]) {
    float4 T = t0m_u(e, pr, u0);
    out[CF_E] = e.x; out[CF_VX] = e.y; out[CF_VY] = e.z; out[CF_VZ] = e.w;
    out[CF_U0] = u0; out[CF_P] = pr;
    out[CF_Q0] = tau * (T.x + pi[0]);
    out[CF_Q1] = tau * (T.y + pi[1]);
    out[CF_Q2] = tau * (T.z + pi[2]);
    out[CF_Q3] = tau * (T.w + pi[3]);
#pragma unroll
    for (int m = 0; m < 10; m++) out[CF_PI0 + m] = pi[m];
    out[CF_BK] = bk;
}

// per-cell quantities of one cell from its raw fields
template <class EOS>
__device__ __forceinline__ void cell_prims(const PhysP &p, const EOS &eos, float tau, float4 e, const float pi[10],
                                           float bk, float out[CF_N]) {
    cell_prims_up(tau, e, pi, bk, gamma3(e.y, e.z, e.w), eos.P(e.x) + bk, out);
}

// Face value between cells L and R along axis A, from the cells Lm = L-1, L, R, Rp = R+1.
// G(field, cell) returns a per-cell quantity.  out: 4 KT components, 10 IS components, Pi.
// HAVE_DAL: the 15 slopes of cell L are given in dal (Lm is not read); WANT_DAR: the 15
// slopes of cell R are returned in dar.  A slope is a function of three consecutive
// cells only, so a slope handed from one face to the next is the value the next face
// would compute.
template <int A, bool BULK, bool HAVE_DAL = false, bool WANT_DAR = false, class EOS, typename G>
__device__ __forceinline__ void face_flux(const PhysP &p, const EOS &eos, float tau, const G &g, int Lm, int L,
                                          int R, int Rp, float out[15], const float *dal = nullptr,
                                          float *dar = nullptr) {
    const float4 eL = make_float4(g(CF_E, L), g(CF_VX, L), g(CF_VY, L), g(CF_VZ, L));
    const float4 eR = make_float4(g(CF_E, R), g(CF_VX, R), g(CF_VY, R), g(CF_VZ, R));
    const float vh = 0.5f * (comp(eL, A + 1) + comp(eR, A + 1));
    const float lam = max_prop_speed(p, eos, half4(eL, eR), vh);
    // KT flux of Q (kernel_visc.cl:86-101)
    const float pr_half = 0.5f * (g(CF_P, L) + g(CF_P, R));
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float DAL = HAVE_DAL ? dal[q] : slope(g(CF_Q0 + q, Lm), g(CF_Q0 + q, L), g(CF_Q0 + q, R));
        float DAR = slope(g(CF_Q0 + q, L), g(CF_Q0 + q, R), g(CF_Q0 + q, Rp));
        if (WANT_DAR) dar[q] = DAR;
        float AL = g(CF_Q0 + q, L) + 0.5f * DAL;
        float AR = g(CF_Q0 + q, R) - 0.5f * DAR;
        const float g0 = q == 0 ? 1.0f : 0.0f;
        const float ga = q == A + 1 ? -1.0f : 0.0f;
        float p0h = 0.5f * (g(CF_PI0 + PI_ROW[0][q], L) + g(CF_PI0 + PI_ROW[0][q], R));
        float pih = 0.5f * (g(CF_PI0 + PI_ROW[A + 1][q], L) + g(CF_PI0 + PI_ROW[A + 1][q], R));
        float Jp = (AR + pr_half * tau * g0 - tau * p0h) * vh - pr_half * tau * ga + tau * pih;
        float Jm = (AL + pr_half * tau * g0 - tau * p0h) * vh - pr_half * tau * ga + tau * pih;
        out[q] = 0.5f * (Jp + Jm) - 0.5f * lam * (AR - AL);
    }
    // IS flux of u^tau pi^{mu nu} (helper.h:275-289)
    const float u0L = g(CF_U0, L), u0R = g(CF_U0, R), u0Rp = g(CF_U0, Rp);
#pragma unroll
    for (int mn = 0; mn < 11; mn++) {
        if (mn == 10 && !BULK) { out[14] = 0.0f; if (WANT_DAR) dar[14] = 0.0f; break; }
        const int f = mn < 10 ? CF_PI0 + mn : CF_BK;
        float QL = u0L * g(f, L), QR = u0R * g(f, R), QRp = u0Rp * g(f, Rp);
        float DAL = HAVE_DAL ? dal[4 + mn] : slope(g(CF_U0, Lm) * g(f, Lm), QL, QR);
        float DAR = slope(QL, QR, QRp);
        if (WANT_DAR) dar[4 + mn] = DAR;
        float AL = QL + 0.5f * DAL;
        float AR = QR - 0.5f * DAR;
        float Jp = AR * vh, Jm = AL * vh;
        out[4 + mn] = 0.5f * (Jp + Jm) - 0.5f * lam * (AR - AL);
    }
}

// A cell whose root solve runs in FP64 (CV_FLAG_FP64_ROOT, M > fp64_vmin T^tt), recorded by
// the stage kernel after the |T^ti| <= 0.999 T^tt clamp and finished by k_fp64_fix.
struct FixEntry {
    uint32_t cell;
    float T00, T01, T02, T03, M, bulk, pad;
};

// Field pointers of one RK stage (the buffers of CLVisc.IS_stepUpdate / visc_stepUpdate).
struct StagePtrs {
    const float *ev_st, *pi_st, *bk_st;        // d_ev[step], d_pi[step], Pi[step]: the stencil
    const float *ev1, *ev2, *pi1, *bk1;        // d_ev[1], d_ev[2], d_pi[1], Pi[1]
    const float *udiff_in;
    float *IS_src, *bsrc, *Src, *pi_out, *bk_out, *ev_out, *good, *udiff_out;
    // split root solve (SPEC & CV_SPEC_SPLIT kernels only): when set, the stage writes the root solve's
    // input (T^tt, T^tx, T^ty, T^t eta before the |T^ti| <= 0.999 T^tt clamp) to four planes
    // here instead of solving; k_visc_root finishes the cell (ev_out, udiff_out)
    float *cons_out;
    // deferred FP64 root solves (SPEC & CV_SPEC_FP64 kernels): list of n entries and its count
    FixEntry *fix;
    unsigned *fix_ctr;
    // learned seed in the stage (SPEC & CV_SPEC_SEED kernels): the network's FP32 parameters
    // and packed operands (k_mlp_pack_t); when set, each cell's seed and its root-solve
    // iteration count (0 for a deferred FP64 cell) are written to seed_out / iters_out
    const float *mlp_P;
    const uint32_t *mlp_pk;
    float *seed_out;
    int *iters_out;
};

// Optional branches of the stage kernels (template mask SPEC; flags and pointers are still
// tested at run time).  Each costs registers in the whole kernel, so a kernel carries only
// what its caller asks for.
#define CV_SPEC_TESTS 1      // GUBSER_VISC_TEST / RIEMANN_TEST / the vorticity couplings
#define CV_SPEC_FP64 2       // FP64 root solve for fast cells
#define CV_SPEC_SPLIT 4      // split root solve (StagePtrs::cons_out)
#define CV_SPEC_ALL 7
#define CV_SPEC_SEED 8       // learned seed of the root solve, run by the whole wave in the stage

// The end of update_ev from the solved e on (kernel_visc.cl:378-392): velocities, the new
// ev and, at STEP 2, udiff = u(ev_new) - u(ev_1) (kernel_IS.cl:714-724).
template <bool BULK, int STEP, class EOS>
__device__ __forceinline__ void update_ev_tail(const EOS &eos, float ed_find, float T00, float T01, float T02,
                                               float T03, float bout, const float u_old[4], float *ev_out,
                                               float *udiff_out, uint32_t bo, uint32_t nb) {
    if (ed_find < CV_ACU) { T01 = 0.0f; T02 = 0.0f; T03 = 0.0f; }
    ed_find = rmax(0.0f, ed_find);
    float pr = eos.P(ed_find) + bout;
    if (BULK) pr = rmax(pr, 0.0f);
    float repv = rcpn(rmax(CV_ACU, T00 + pr));         // T00 + pr >= 1e-7 (rcpn's range)
    float4 en = make_float4(ed_find, T01 * repv, T02 * repv, T03 * repv);
    stp(ev_out, bo, nb, 0, en.x);
    stp(ev_out, bo, nb, 1, en.y);
    stp(ev_out, bo, nb, 2, en.z);
    stp(ev_out, bo, nb, 3, en.w);
    if (STEP == 2) {
        float gn = gamma3(en.y, en.z, en.w);
        float un[4] = {gn * 1.0f, gn * en.y, gn * en.z, gn * en.w};
#pragma unroll
        for (int c = 0; c < 4; c++) stp(udiff_out, bo, nb, c, un[c] - u_old[c]);
    }
}

// One work item of the stage.  All threads of the workgroup call it (it synchronises).
// EK: EOS table kind (0 row/col, 1 octave); SPEC: mask of the optional branches above.
// MARCH = false: a work item is one tile of one x-plane; every thread evaluates its two x
// faces from its own x-stencil (cells i-2..i+2, read from global memory).
// MARCH = true: a work item is one tile column over a chunk of xc consecutive x-planes; the
// workgroup walks the chunk in increasing x and each thread carries the x face F(i-1/2),
// the slopes of its cell i and u^mu of cell i-1 from one plane to the next, so every x face
// and every x slope is evaluated once (the chunk's first plane starts from the full stencil).
template <bool BULK, int STEP, int EK, int SPEC, bool MARCH, class EOS>
__device__ __forceinline__ void tile_work(const PhysP &p, const EOS &eos, const StageP &st, const TileP &tp,
                                          const StagePtrs &q_, int64_t work, float (*s_c)[CV_TILE_MAXC],
                                          float (*s_f)[CV_TILE_MAXF]) {
    const float *__restrict__ ev_st = q_.ev_st;
    const float *__restrict__ pi_st = q_.pi_st;
    const float *__restrict__ bk_st = q_.bk_st;
    const float *ev1 = q_.ev1, *ev2 = q_.ev2, *pi1 = q_.pi1, *bk1 = q_.bk1;
    const float *__restrict__ udiff_in = q_.udiff_in;
    float *IS_src = q_.IS_src, *bsrc = q_.bsrc, *Src = q_.Src, *pi_out = q_.pi_out, *bk_out = q_.bk_out;
    float *ev_out = q_.ev_out, *good = q_.good, *udiff_out = q_.udiff_out;
    const int tid = threadIdx.x;
    const uint32_t nb = p.nb;
    const float tau = st.tau;
    const int ty = tp.ty, tz = tp.tz, HY = tp.HY, HZ = tp.HZ, hy = tp.hy, hz = tp.hz;
    const int my = tid / tz, mz = tid - (tid / tz) * tz;       // own cell in the tile
    const bool lane_in_tile = tid < ty * tz;
    {
        const int kt = (int)(work % tp.ntk);
        const int64_t r = work / tp.ntk;
        const int jt = (int)(r % tp.ntj);
        const int xr = (int)(r / tp.ntj);            // MARCH: chunk, else the x-plane
        int ibeg = xr, iend = xr + 1;
        if (MARCH) march_planes(tp, p.nx, xr, ibeg, iend);
        if (ibeg >= iend) return;                    // no planes (uniform over the workgroup)
        const int j0 = jt * ty, k0 = kt * tz;
        // state carried along x by MARCH: F(i-1/2), slopes of cell i, u^mu of cell i-1, and
        // u^tau, P + Pi of cell i+1 (evaluated as cell i+2 of the previous plane)
        float xFm[15], xDAL[15], xum1[4], xu0n = 0.0f, xprn = 0.0f;
        if (MARCH && lane_in_tile && j0 + my < p.ny && k0 + mz < p.nz) {
            const uint32_t I0 = (uint32_t)((int64_t)ibeg * p.sx) + (uint32_t)(j0 + my) * (uint32_t)p.sy +
                                (uint32_t)(k0 + mz);
            float xc[4][CF_N];                                      // cells ibeg-2 .. ibeg+1
#pragma unroll
            for (int w = 0; w < 4; w++) {
                int ii = ibeg + w - 2;
                ii = ii < 0 ? 0 : (ii > p.nx - 1 ? p.nx - 1 : ii);
                const uint32_t b2 = (I0 + (uint32_t)(ii - ibeg) * (uint32_t)p.sx) * 4u;
                float4 e = ld4p(ev_st, b2, nb);
                float pv[10];
#pragma unroll
                for (int m = 0; m < 10; m++) pv[m] = ldp(pi_st, b2, nb, m);
                float bkv = BULK ? ldo(bk_st, b2) : 0.0f;
                cell_prims(p, eos, tau, e, pv, bkv, xc[w]);
            }
            auto gx = [&](int f, int c) { return xc[c][f]; };
            face_flux<0, BULK, false, true>(p, eos, tau, gx, 0, 1, 2, 3, xFm, nullptr, xDAL);
#pragma unroll
            for (int c = 0; c < 4; c++) xum1[c] = c == 0 ? xc[1][CF_U0] * 1.0f : xc[1][CF_U0] * xc[1][CF_E + c];
            xu0n = xc[3][CF_U0];
            xprn = xc[3][CF_P];
        }
      for (int i = ibeg; i < iend; i++) {
        const uint32_t plane = (uint32_t)((int64_t)i * p.sx);

        // ---- 1. per-cell quantities of tile + y/z halo into LDS ----
        for (int c = tid; c < HY * HZ; c += CV_TILE_THREADS) {
            const int ly = c / HZ, lz = c - (c / HZ) * HZ;
            const bool hy_out = ly < hy || ly >= hy + ty, hz_out = lz < hz || lz >= hz + tz;
            if (hy_out && hz_out) continue;                       // corner: not needed
            int j = j0 + ly - hy, k = k0 + lz - hz;
            j = j < 0 ? 0 : (j > p.ny - 1 ? p.ny - 1 : j);
            k = k < 0 ? 0 : (k > p.nz - 1 ? p.nz - 1 : k);
            const uint32_t bo = (plane + (uint32_t)j * (uint32_t)p.sy + (uint32_t)k) * 4u;
            float4 e = ld4p(ev_st, bo, nb);
            float pv[10];
#pragma unroll
            for (int m = 0; m < 10; m++) pv[m] = ldp(pi_st, bo, nb, m);
            float bkv = BULK ? ldo(bk_st, bo) : 0.0f;
            float out[CF_N];
            cell_prims(p, eos, tau, e, pv, bkv, out);
#pragma unroll
            for (int f = 0; f < CF_N; f++) s_c[f][c] = out[f];
        }
        __syncthreads();

        const int j = j0 + my, k = k0 + mz;
        const bool active = lane_in_tile && j < p.ny && k < p.nz;
        const int cc = (my + hy) * HZ + (mz + hz);                  // own LDS cell
        const uint32_t I = plane + (uint32_t)j * (uint32_t)p.sy + (uint32_t)k;
        const uint32_t bo = I * 4u;
        auto gl = [&](int f, int c) { return s_c[f][c]; };

        float s[10], bs = 0.0f, t[4], udx[4], udy[4], udz[4];
        float q[10];
        float4 ec = make_float4(0.f, 0.f, 0.f, 0.f);
        float T_pre = 0.0f, s_pre = 0.0f, e2x_pre = 0.0f, u0_own = 1.0f, bk_own = 0.0f;
        if (active) {
            if (STEP == 1) {
#pragma unroll
                for (int m = 0; m < 10; m++) s[m] = 0.0f;
                t[0] = t[1] = t[2] = t[3] = 0.0f;
            } else {
#pragma unroll
                for (int m = 0; m < 10; m++) s[m] = ldp(IS_src, bo, nb, m);
                if (BULK) bs = ldo(bsrc, bo);
                float4 qq = ld4p(Src, bo, nb);
                t[0] = qq.x; t[1] = qq.y; t[2] = qq.z; t[3] = qq.w;
            }
            ec = make_float4(s_c[CF_E][cc], s_c[CF_VX][cc], s_c[CF_VY][cc], s_c[CF_VZ][cc]);
            u0_own = s_c[CF_U0][cc];
            if (BULK) bk_own = s_c[CF_BK][cc];
            // the update's EOS lookup (T, s at the stage's own e) and, at STEP 1, the stale e
            // of d_ev[2] are issued here, so that their latency overlaps the face passes
            {
                const float4 cpTs0 = eos.raw(ec.x);
                T_pre = cpTs0.z;
                s_pre = cpTs0.w;
            }
            if (STEP == 1) e2x_pre = ldo(ev2, bo);
#pragma unroll
            for (int m = 0; m < 10; m++) q[m] = s_c[CF_PI0 + m][cc];
            // visc_src_christoffel (kernel_IS.cl:41-67)
            {
                float uz = ec.w * s_c[CF_U0][cc];
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
                float u0 = s_c[CF_U0][cc];
                float pressure = s_c[CF_P][cc];
                float Tzz = (ec.x + pressure) * u0 * u0 * ec.w * ec.w + pressure + q[9];
                float Ttz = (ec.x + pressure) * u0 * u0 * ec.w + q[3];
                t[0] = t[0] - Tzz; t[1] = t[1] - 0.0f; t[2] = t[2] - 0.0f; t[3] = t[3] - Ttz;
                if ((SPEC & CV_SPEC_TESTS) && (p.flags & 1)) {
                    float Ttt = (ec.x + pressure) * u0 * u0 - pressure;
                    float Ttx = (ec.x + pressure) * u0 * u0 * ec.y;
                    float Tty = (ec.x + pressure) * u0 * u0 * ec.z;
                    t[0] = t[0] + (Tzz + Ttt); t[1] = t[1] + Ttx; t[2] = t[2] + Tty; t[3] = t[3] + 3.0f * Ttz;
                }
            }
            // ---- 2. x ----
            if (MARCH) {
                // cells i+1 and i+2 of the own column; cell i is in LDS
                float xr1[CF_N], xr2[CF_N];
#pragma unroll
                for (int w = 1; w <= 2; w++) {
                    int ii = i + w;
                    ii = ii > p.nx - 1 ? p.nx - 1 : ii;
                    const uint32_t b2 = (I + (uint32_t)(ii - i) * (uint32_t)p.sx) * 4u;
                    float4 e = ld4p(ev_st, b2, nb);
                    float pv[10];
#pragma unroll
                    for (int m = 0; m < 10; m++) pv[m] = ldp(pi_st, b2, nb, m);
                    float bkv = BULK ? ldo(bk_st, b2) : 0.0f;
                    if (w == 1) {
                        // cell i+1 (min(i+1, nx-1)) is the previous plane's cell i+2
                        // (min(i+2, nx-1) there), with the same stage inputs
                        cell_prims_up(tau, e, pv, bkv, xu0n, xprn, xr1);
                    } else {
                        cell_prims(p, eos, tau, e, pv, bkv, xr2);
                        xu0n = xr2[CF_U0];
                        xprn = xr2[CF_P];
                    }
                }
                auto gx = [&](int f, int c) { return c == 1 ? s_c[f][cc] : (c == 2 ? xr1[f] : xr2[f]); };
                float Fp[15], dar[15];
                face_flux<0, BULK, true, true>(p, eos, tau, gx, 0, 1, 2, 3, Fp, xDAL, dar);
#pragma unroll
                for (int m = 0; m < 4; m++) t[m] = t[m] - (Fp[m] - xFm[m]) * p.r_dx;
#pragma unroll
                for (int m = 0; m < 10; m++) s[m] = s[m] - (Fp[4 + m] - xFm[4 + m]) * p.r_dx;
                if (BULK) bs = bs - (Fp[14] - xFm[14]) * p.r_dx;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    float ui = c == 0 ? s_c[CF_U0][cc] * 1.0f : s_c[CF_U0][cc] * s_c[CF_E + c][cc];
                    float up1 = c == 0 ? xr1[CF_U0] * 1.0f : xr1[CF_U0] * xr1[CF_E + c];
                    udx[c] = slope(xum1[c], ui, up1) * p.r_dx;
                    xum1[c] = ui;
                }
#pragma unroll
                for (int m = 0; m < 15; m++) { xFm[m] = Fp[m]; xDAL[m] = dar[m]; }
            } else {
                float xc[5][CF_N];
#pragma unroll
                for (int w = 0; w < 5; w++) {
                    if (w == 2) {
#pragma unroll
                        for (int f = 0; f < CF_N; f++) xc[2][f] = s_c[f][cc];
                    } else {
                        int ii = i + w - 2;
                        ii = ii < 0 ? 0 : (ii > p.nx - 1 ? p.nx - 1 : ii);
                        const uint32_t b2 = (I + (uint32_t)(ii - i) * (uint32_t)p.sx) * 4u;
                        float4 e = ld4p(ev_st, b2, nb);
                        float pv[10];
#pragma unroll
                        for (int m = 0; m < 10; m++) pv[m] = ldp(pi_st, b2, nb, m);
                        float bkv = BULK ? ldo(bk_st, b2) : 0.0f;
                        cell_prims(p, eos, tau, e, pv, bkv, xc[w]);
                    }
                }
                auto gx = [&](int f, int c) { return xc[c][f]; };
                float Fp[15], Fm[15];
                face_flux<0, BULK>(p, eos, tau, gx, 1, 2, 3, 4, Fp);
                face_flux<0, BULK>(p, eos, tau, gx, 0, 1, 2, 3, Fm);
#pragma unroll
                for (int m = 0; m < 4; m++) t[m] = t[m] - (Fp[m] - Fm[m]) * p.r_dx;
#pragma unroll
                for (int m = 0; m < 10; m++) s[m] = s[m] - (Fp[4 + m] - Fm[4 + m]) * p.r_dx;
                if (BULK) bs = bs - (Fp[14] - Fm[14]) * p.r_dx;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    float um1 = c == 0 ? xc[1][CF_U0] * 1.0f : xc[1][CF_U0] * xc[1][CF_E + c];
                    float ui = c == 0 ? xc[2][CF_U0] * 1.0f : xc[2][CF_U0] * xc[2][CF_E + c];
                    float up1 = c == 0 ? xc[3][CF_U0] * 1.0f : xc[3][CF_U0] * xc[3][CF_E + c];
                    udx[c] = slope(um1, ui, up1) * p.r_dx;
                }
            }
        }

        // ---- 3. y faces: face f in [0, ty] lies between tile rows f-1 and f ----
        if (tp.hy > 0) {
            for (int fc = tid; fc < (ty + 1) * tz; fc += CV_TILE_THREADS) {
                const int f = fc / tz, z = fc - (fc / tz) * tz;
                const int Lc = (f - 1 + hy) * HZ + (z + hz);
                float F[15];
                face_flux<1, BULK>(p, eos, tau, gl, Lc - HZ, Lc, Lc + HZ, Lc + 2 * HZ, F);
#pragma unroll
                for (int m = 0; m < 15; m++) s_f[m][fc] = F[m];
            }
        }
        __syncthreads();
        if (active) {
            if (tp.hy > 0) {
                const int fp_ = (my + 1) * tz + mz, fm_ = my * tz + mz;
#pragma unroll
                for (int m = 0; m < 4; m++) t[m] = t[m] - (s_f[m][fp_] - s_f[m][fm_]) * p.r_dy;
#pragma unroll
                for (int m = 0; m < 10; m++) s[m] = s[m] - (s_f[4 + m][fp_] - s_f[4 + m][fm_]) * p.r_dy;
                if (BULK) bs = bs - (s_f[14][fp_] - s_f[14][fm_]) * p.r_dy;
            } else {
                // ny == 1: both faces are the same expression of the same cells -> 0
#pragma unroll
                for (int m = 0; m < 4; m++) t[m] = t[m] - 0.0f * p.r_dy;
#pragma unroll
                for (int m = 0; m < 10; m++) s[m] = s[m] - 0.0f * p.r_dy;
                if (BULK) bs = bs - 0.0f * p.r_dy;
            }
#pragma unroll
            for (int c = 0; c < 4; c++) {
                float um1, ui, up1;
                const int cm = tp.hy > 0 ? cc - HZ : cc, cp = tp.hy > 0 ? cc + HZ : cc;
                um1 = c == 0 ? s_c[CF_U0][cm] * 1.0f : s_c[CF_U0][cm] * s_c[CF_E + c][cm];
                ui = c == 0 ? s_c[CF_U0][cc] * 1.0f : s_c[CF_U0][cc] * s_c[CF_E + c][cc];
                up1 = c == 0 ? s_c[CF_U0][cp] * 1.0f : s_c[CF_U0][cp] * s_c[CF_E + c][cp];
                udy[c] = slope(um1, ui, up1) * p.r_dy;
            }
            // z gradient of u^mu (cells only) here, so that no cell value is read after the
            // z faces: the next plane may then overwrite the cells without a barrier of its own
            const float u0c = u0_own;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int cm = tp.hz > 0 ? cc - 1 : cc, cp = tp.hz > 0 ? cc + 1 : cc;
                float um1 = c == 0 ? s_c[CF_U0][cm] * 1.0f : s_c[CF_U0][cm] * s_c[CF_E + c][cm];
                float ui = c == 0 ? u0c * 1.0f : u0c * s_c[CF_E + c][cc];
                float up1 = c == 0 ? s_c[CF_U0][cp] * 1.0f : s_c[CF_U0][cp] * s_c[CF_E + c][cp];
                float chr = ((c == 0 ? ec.w : (c == 3 ? 1.0f : 0.0f)) * u0c) * st.r_tau;
                udz[c] = slope(um1, ui, up1) * st.r_taudz + chr;
            }
        }
        __syncthreads();
        // ---- 4. z faces: face f in [0, tz] lies between tile columns f-1 and f ----
        if (tp.hz > 0) {
            for (int fc = tid; fc < ty * (tz + 1); fc += CV_TILE_THREADS) {
                const int y = fc / (tz + 1), f = fc - (fc / (tz + 1)) * (tz + 1);
                const int Lc = (y + hy) * HZ + (f - 1 + hz);
                float F[15];
                face_flux<2, BULK>(p, eos, tau, gl, Lc - 1, Lc, Lc + 1, Lc + 2, F);
#pragma unroll
                for (int m = 0; m < 15; m++) s_f[m][fc] = F[m];
            }
        }
        __syncthreads();
        // the update's reads (STEP 1: udiff; STEP 2: d_ev[1], d_pi[1] and Pi[1]) are issued
        // here, ahead of the z differences, so that their latency overlaps that work
        float4 e1_pre = ec;
        float pi1_pre[10], bk1_pre = 0.0f, ud_pre[4];
        if (STEP == 1 && active) {
#pragma unroll
            for (int c = 0; c < 4; c++) ud_pre[c] = ldp(udiff_in, bo, nb, c);
        }
        if (STEP == 2 && active) {
            e1_pre = ld4p(ev1, bo, nb);
#pragma unroll
            for (int m = 0; m < 10; m++) pi1_pre[m] = ldp(pi1, bo, nb, m);
            if (BULK) bk1_pre = ldo(bk1, bo);
        }
        if (active) {
            if (tp.hz > 0) {
                const int fp_ = my * (tz + 1) + mz + 1, fm_ = my * (tz + 1) + mz;
#pragma unroll
                for (int m = 0; m < 4; m++) t[m] = t[m] - (s_f[m][fp_] - s_f[m][fm_]) * st.r_taudz;
#pragma unroll
                for (int m = 0; m < 10; m++) s[m] = s[m] - (s_f[4 + m][fp_] - s_f[4 + m][fm_]) * st.r_taudz;
                if (BULK) bs = bs - (s_f[14][fp_] - s_f[14][fm_]) * st.r_taudz;
            } else {
#pragma unroll
                for (int m = 0; m < 4; m++) t[m] = t[m] - 0.0f * st.r_taudz;
#pragma unroll
                for (int m = 0; m < 10; m++) s[m] = s[m] - 0.0f * st.r_taudz;
                if (BULK) bs = bs - 0.0f * st.r_taudz;
            }
        }

        // ---- 5. update_pimn (kernel_IS.cl:501-709) and update_ev (kernel_visc.cl:322-393) ----
        // SEED: the root-solve input of the cell, carried out of the per-lane branch so that
        // the whole wave runs the network (sd_go: the cell solves from the seed)
        float sd_T[4] = {0.0f, 0.0f, 0.0f, 0.0f}, sd_M = 0.0f, sd_b = 0.0f;
        float sd_u[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        uint32_t sd_lo = 0u, sd_hi = 0u;
        float sd_t00 = 0.0f;                                   // T^tt of the cell, for the seed floor
        bool sd_go = false;
        if (active) {
            const float4 e1 = STEP == 1 ? ec : e1_pre;
            const float4 e2 = STEP == 2 ? ec : make_float4(e2x_pre, 0.0f, 0.0f, 0.0f);
            float g1 = STEP == 1 ? u0_own : gamma3(e1.y, e1.z, e1.w);
            float u_old[4] = {g1 * 1.0f, g1 * e1.y, g1 * e1.z, g1 * e1.w};
            float u_new[4], udt[4];
            if (STEP == 1) {
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    float ud_ = ud_pre[c];
                    udt[c] = ud_ * p.r_dt;
                    u_new[c] = u_old[c] + ud_;
                }
            } else {
                float g2 = u0_own;
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
            const float4 cpTs = make_float4(0.0f, 0.0f, T_pre, s_pre);   // eos.raw(ed_step), ed_step = ec.x
            const float temperature = cpTs.z;
            const float local_etaos = etaos(p, temperature);
            float etav = local_etaos * cpTs.w * CV_HBARC;
            // rcpn: the operands below are >= 5.9e-8 (acu floor) and >= 1 (u^tau + ...)
            float one_over_taupi = temperature * rcpn(3.0f * rmax(CV_ACU, local_etaos) * CV_HBARC);
            float coef_pipi = 0.0f;
            const bool gubser = (SPEC & CV_SPEC_TESTS) && (p.flags & 2);
            // vorticity couplings (test-branch instances only; production kernels carry neither)
            const bool cpl = (SPEC & CV_SPEC_TESTS) && (p.flags & (CV_FLAG_PIMN_OMEGA | CV_FLAG_OMEGA_OMEGA));
            float Om[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            if (cpl) omega_cell(udt, udx, udy, udz, u, Om);
            if (gubser) {
                float q4 = sqrtf(sqrtf(ed_step));
                float ETAOS = p.etaos_ymin;
                float etavH = (ETAOS * 4.0f) * p.r_3;
                float taupiH = (p.lam1 * p.lam1 * etavH) * p.r_3;
                one_over_taupi = rcp(taupiH * rcp(q4));
                coef_pipi = p.lam1 * rcp(ed_step);
                etav = ((ETAOS * 4.0f) * p.r_3) * (sqrtf(ed_step) * q4);
            }
            const float gd[4] = {1.0f, -1.0f, -1.0f, -1.0f};
            float max_abs = 0.0f;
            float out[10];
            const float rden = rcpn(u_new[0] + p.dt * one_over_taupi);
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
                    float pi_old = (STEP == 1 ? q[mn] : pi1_pre[mn]) * u_old[0];
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
                float oot_b = temperature * rcpn(3.0f * rmax(CV_ACU, zs) * CV_HBARC);
                float b2 = bk_own;
                float bsrc_ = (b2 * theta) * p.r_3 - (b2 * u[0]) * st.r_tau;
                bs += bsrc_;
                b_start = STEP == 1 ? b2 : bk1_pre;
                float b_old = b_start * u_old[0];
                bout = (b_old + (bs * p.dt) * st.r_step - p.dt * oot_b * zetav * theta) * rcpn(u_new[0] + p.dt * oot_b);
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
                *(float *)((char *)good + bo) = gc;
            }
#pragma unroll
            for (int m = 0; m < 10; m++) stp(pi_out, bo, nb, m, out[m]);
            if (BULK) *(float *)((char *)bk_out + bo) = bout;
            if (STEP == 1) {
#pragma unroll
                for (int m = 0; m < 10; m++) stp(IS_src, bo, nb, m, s[m]);
                if (BULK) *(float *)((char *)bsrc + bo) = bs;
#pragma unroll
                for (int m = 0; m < 4; m++) stp(Src, bo, nb, m, t[m]);
            }
            // update_ev
            float pressure = eos.P(e1.x) + (BULK ? b_start : 0.0f);
            float T0m[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                float g0 = c == 0 ? 1.0f : 0.0f;
                float pim0 = STEP == 1 ? q[c] : pi1_pre[c];
                T0m[c] = ((e1.x + pressure) * g1 * u_old[c] - pressure * g0 + pim0) * st.old_time;
            }
#pragma unroll
            for (int c = 0; c < 4; c++) T0m[c] = T0m[c] + (t[c] * p.dt) * st.r_step - out[c] * st.tau;
            float T00 = rmax(CV_ACU, T0m[0]) * st.r_tau;
            float T01 = fabsf(T0m[1]) < CV_ACU ? 0.0f : T0m[1] * st.r_tau;
            float T02 = fabsf(T0m[2]) < CV_ACU ? 0.0f : T0m[2] * st.r_tau;
            float T03 = fabsf(T0m[3]) < CV_ACU ? 0.0f : T0m[3] * st.r_tau;
            if ((SPEC & CV_SPEC_SPLIT) && q_.cons_out) {
                stp(q_.cons_out, bo, nb, 0, T00);
                stp(q_.cons_out, bo, nb, 1, T01);
                stp(q_.cons_out, bo, nb, 2, T02);
                stp(q_.cons_out, bo, nb, 3, T03);
            } else {
            float M = sqrtf(T01 * T01 + T02 * T02 + T03 * T03);
            if (M > T00) {
                float sc = (0.999f * T00) * rcpn(M);            // M > T00 >= 1e-7 / tau
                T01 *= sc; T02 *= sc; T03 *= sc;
                M = 0.999f * T00;
            }
            if (SPEC & CV_SPEC_SEED) { seed_features_bf16(T00, M, T03, sd_lo, sd_hi); sd_t00 = T00; }
            if ((SPEC & CV_SPEC_FP64) && (p.flags & 8) && M > p.fp64_vmin * T00) {
                // FP64 root solve deferred to k_fp64_fix (kept out of this kernel's registers);
                // each cell is recorded at most once per stage, so n entries always suffice.
                // One atomic per wavefront: the deferring lanes take consecutive slots.
                const uint64_t mask = __ballot(1);
                const int lane = (int)__lane_id();
                const int leader = __ffsll((unsigned long long)mask) - 1;
                unsigned base = 0u;
                if (lane == leader) base = atomicAdd(q_.fix_ctr, (unsigned)__popcll(mask));
                base = __shfl(base, leader, 64);
                const unsigned slot = base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
                q_.fix[slot] = FixEntry{bo >> 2, T00, T01, T02, T03, M, bout, 0.0f};
                if ((SPEC & CV_SPEC_SEED) && q_.iters_out) q_.iters_out[bo >> 2] = 0;
            } else if (SPEC & CV_SPEC_SEED) {
                sd_go = true;
                sd_T[0] = T00; sd_T[1] = T01; sd_T[2] = T02; sd_T[3] = T03;
                sd_M = M;
                sd_b = bout;
#pragma unroll
                for (int c = 0; c < 4; c++) sd_u[c] = u_old[c];
            } else {
                update_ev_tail<BULK, STEP>(eos, root_newton(eos, T00, M, bout, BULK), T00, T01, T02, T03, bout, u_old,
                                           ev_out, udiff_out, bo, nb);
            }
            }
        }
        if constexpr ((SPEC & CV_SPEC_SEED) != 0) {
            // the network on the wave's 64 cells (lanes without a cell give zero features) when
            // one of them is above the seed floor, then the root solve from the seed with the
            // reference's bracket and stopping rule
            const bool wave_net = __any(sd_t00 >= MLP_SEED_TMIN);
            const float v0 = seed_gate(sd_t00, wave_net ? seed_wave_t(q_.mlp_P, q_.mlp_pk, sd_lo, sd_hi) : 0.5f);
            if (active && q_.seed_out) *(float *)((char *)q_.seed_out + bo) = v0;
            if (sd_go) {
                int it = 0;
                const float ed_find = root_newton_from(eos, sd_T[0], sd_M, sd_b, BULK, fminf(fmaxf(v0, 0.0f), 0.9999f), &it);
                if (q_.iters_out) q_.iters_out[bo >> 2] = it;
                update_ev_tail<BULK, STEP>(eos, ed_find, sd_T[0], sd_T[1], sd_T[2], sd_T[3], sd_b, sd_u, ev_out,
                                           udiff_out, bo, nb);
            }
        }
        // no barrier here: after the z faces no cell value is read, and the next plane's
        // face passes (which overwrite the faces) come after that plane's first barrier
      }
    }
}

template <bool BULK, int STEP, int EK, int SPEC, bool MARCH>
__global__ __launch_bounds__(CV_TILE_THREADS, MARCH ? 3 : 2) void k_visc_tiled(PhysP p, EosP E, StageP st, TileP tp,
                                                                              StagePtrs q) {
    __shared__ float s_c[CF_N][CV_TILE_MAXC];
    __shared__ float s_f[15][CV_TILE_MAXF];
    const int tid = threadIdx.x;
    const int64_t total = MARCH ? tp.nwork : tp.ntiles;
    if constexpr (MARCH && EK == 1) {
        // the pressure column of the octave table (the host launches this kernel only for
        // tables of at most CV_PTAB_LDS nodes)
        __shared__ float s_pt[CV_PTAB_LDS];
        for (int k = tid; k <= E.nlast; k += CV_TILE_THREADS) s_pt[k] = E.tab[k].y;
        __syncthreads();
        EosOctP eos(E, E.tab, s_pt);
        if (tp.xcd > 0 && (int64_t)gridDim.x == total) {
            tile_work<BULK, STEP, EK, SPEC, MARCH>(p, eos, st, tp, q, xcd_work(blockIdx.x, total, tp.xcd), s_c, s_f);
            return;
        }
        for (int64_t work = blockIdx.x; work < total; work += gridDim.x)
            tile_work<BULK, STEP, EK, SPEC, MARCH>(p, eos, st, tp, q, work, s_c, s_f);
    } else {
        // EOS table into LDS (octave table small enough), else global
        __shared__ float4 s_eos[(!MARCH && EK == 1) ? CV_EOS_LDS : 1];
        const float4 *tab = E.tab;
        if (CV_TILE_EOS_IN_LDS && !MARCH && EK == 1 && E.nlast + 1 <= CV_EOS_LDS) {
            for (int k = tid; k <= E.nlast; k += CV_TILE_THREADS) s_eos[k] = E.tab[k];
            tab = s_eos;
        }
        __syncthreads();
        EosT<EK> eos(E, tab);
        if (tp.xcd > 0 && (int64_t)gridDim.x == total) {
            tile_work<BULK, STEP, EK, SPEC, MARCH>(p, eos, st, tp, q, xcd_work(blockIdx.x, total, tp.xcd), s_c, s_f);
            return;
        }
        for (int64_t work = blockIdx.x; work < total; work += gridDim.x)
            tile_work<BULK, STEP, EK, SPEC, MARCH>(p, eos, st, tp, q, work, s_c, s_f);
    }
}

// The end of update_ev for the split stage (kernel_visc.cl:322-393 from the clamp on), one
// cell: reads the conserved input the stage wrote to cons (four planes of n), Pi_new (bk_out)
// and, at STEP 2, d_ev[1] for udiff; solves for e in FP64 when the cell is fast
// (CV_FLAG_FP64_ROOT, M > fp64_vmin T^tt), else from seed_v (when given) or from the
// reference's start; writes ev_out, udiff_out (STEP 2) and the iteration count.  The same
// FP32 operations as the fused stage, so the split stage without a seed is bit-identical.
template <bool BULK, int STEP, int EK>
__device__ __forceinline__ void visc_root_cell(const PhysP &p, const EosP &E, int64_t I, const float *__restrict__ cons,
                                               const float *__restrict__ bk_out, const float *__restrict__ ev1,
                                               const float *__restrict__ seed_v, float *ev_out, float *udiff_out,
                                               int *iters_out) {
    const int64_t n = p.n;
    EosT<EK> eos(E, E.tab);
    float T00 = cons[I], T01 = cons[n + I], T02 = cons[2 * n + I], T03 = cons[3 * n + I];
    const float bout = BULK ? bk_out[I] : 0.0f;
    float M = sqrtf(T01 * T01 + T02 * T02 + T03 * T03);
    if (M > T00) {
        float sc = (0.999f * T00) * rcpn(M);                // M > T00 >= 1e-7 / tau
        T01 *= sc; T02 *= sc; T03 *= sc;
        M = 0.999f * T00;
    }
    float ed_find;
    int it = 0;
    if ((p.flags & 8) && M > p.fp64_vmin * T00)
        ed_find = (float)root_newton_d_inl(E, E.tab, (double)T00, (double)M, (double)bout, BULK);
    else if (seed_v)
        ed_find = root_newton_from(eos, T00, M, bout, BULK, fminf(fmaxf(seed_v[I], 0.0f), 0.9999f), &it);
    else
        ed_find = root_newton_from(eos, T00, M, bout, BULK, 0.5f * (0.0f + 1.0f), &it);
    if (iters_out) iters_out[I] = it;
    if (ed_find < CV_ACU) { T01 = 0.0f; T02 = 0.0f; T03 = 0.0f; }
    ed_find = rmax(0.0f, ed_find);
    float pr = eos.P(ed_find) + bout;
    if (BULK) pr = rmax(pr, 0.0f);
    float repv = rcpn(rmax(CV_ACU, T00 + pr));         // T00 + pr >= 1e-7 (rcpn's range)
    float4 en = make_float4(ed_find, T01 * repv, T02 * repv, T03 * repv);
    ev_out[I] = en.x;
    ev_out[n + I] = en.y;
    ev_out[2 * n + I] = en.z;
    ev_out[3 * n + I] = en.w;
    if (STEP == 2) {
        const float v1x = ev1[n + I], v1y = ev1[2 * n + I], v1z = ev1[3 * n + I];
        const float g1 = gamma3(v1x, v1y, v1z);
        const float u_old[4] = {g1 * 1.0f, g1 * v1x, g1 * v1y, g1 * v1z};
        float gn = gamma3(en.y, en.z, en.w);
        float un[4] = {gn * 1.0f, gn * en.y, gn * en.z, gn * en.w};
#pragma unroll
        for (int c = 0; c < 4; c++) udiff_out[c * n + I] = un[c] - u_old[c];
    }
}

template <bool BULK, int STEP, int EK>
__global__ __launch_bounds__(256) void k_visc_root(PhysP p, EosP E, const float *__restrict__ cons,
                                                   const float *__restrict__ bk_out, const float *__restrict__ ev1,
                                                   const float *__restrict__ seed_v, float *ev_out, float *udiff_out,
                                                   int *iters_out) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= p.n) return;
    visc_root_cell<BULK, STEP, EK>(p, E, I, cons, bk_out, ev1, seed_v, ev_out, udiff_out, iters_out);
}

// The end of update_ev for one deferred cell: FP64 root solve, then the same FP32 tail as
// the stage kernel (kernel_visc.cl:375-392), so the fields are those of the inline solve.
template <bool BULK, int STEP, int EK>
__device__ __forceinline__ void fp64_fix_cell(const PhysP &p, const EosP &E, const FixEntry &f,
                                              const float *__restrict__ ev1, float *ev_out, float *udiff_out) {
    EosT<EK> eos(E, E.tab);
    const int64_t n = p.n, I = f.cell;
    float T01 = f.T01, T02 = f.T02, T03 = f.T03;
    float ed_find = (float)root_newton_d_inl(E, E.tab, (double)f.T00, (double)f.M, (double)f.bulk, BULK);
    if (ed_find < CV_ACU) { T01 = 0.0f; T02 = 0.0f; T03 = 0.0f; }
    ed_find = rmax(0.0f, ed_find);
    float pr = eos.P(ed_find) + f.bulk;
    if (BULK) pr = rmax(pr, 0.0f);
    float repv = rcpn(rmax(CV_ACU, f.T00 + pr));
    float4 en = make_float4(ed_find, T01 * repv, T02 * repv, T03 * repv);
    ev_out[I] = en.x;
    ev_out[n + I] = en.y;
    ev_out[2 * n + I] = en.z;
    ev_out[3 * n + I] = en.w;
    if (STEP == 2) {
        const float v1x = ev1[n + I], v1y = ev1[2 * n + I], v1z = ev1[3 * n + I];
        const float g1 = gamma3(v1x, v1y, v1z);
        const float u_old[4] = {g1 * 1.0f, g1 * v1x, g1 * v1y, g1 * v1z};
        float gn = gamma3(en.y, en.z, en.w);
        float un[4] = {gn * 1.0f, gn * en.y, gn * en.z, gn * en.w};
#pragma unroll
        for (int c = 0; c < 4; c++) udiff_out[c * n + I] = un[c] - u_old[c];
    }
}

template <bool BULK, int STEP, int EK>
__global__ __launch_bounds__(256) void k_fp64_fix(PhysP p, EosP E, const FixEntry *__restrict__ fix,
                                                  const unsigned *__restrict__ ctr, const float *__restrict__ ev1,
                                                  float *ev_out, float *udiff_out) {
    const unsigned cnt = *ctr;
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gridDim.x * blockDim.x)
        fp64_fix_cell<BULK, STEP, EK>(p, E, fix[q], ev1, ev_out, udiff_out);
}

__global__ void k_fix_reset(unsigned *ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *ctr = 0u;
}

}  // namespace cv
