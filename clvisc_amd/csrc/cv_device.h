// cv_device.h -- device-side arithmetic of the KT + Israel-Stewart hydro step (gfx950).
//
// Every function evaluates the formulas of the reference kernels
// (pyvisc/kernel/helper.h, eos_table.h, kernel_ideal.cl, kernel_visc.cl,
// kernel_IS.cl) under one arithmetic contract:
//   * a/b is a * RN(1/b); RN(1/b) of a run constant is computed once on the host,
//   * sqrt is correctly rounded, nothing is contracted into FMA (this file is
//     compiled with -ffp-contract=off),
//   * operands are combined in the order the reference writes them.
// Under this contract an IEEE CPU evaluating the same operations gets the same
// bits, which is how the FP32 results are checked.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CV_ACU 1.0e-7f
#define CV_BULK_EMIN 1.0e-3f   // bulk Pi is switched off below this e [GeV/fm^3]
#define CV_THETA 1.1f
#define CV_HBARC 0.19733f

namespace cv {

// ---------------------------------------------------------------------------
// Parameters passed by value to every kernel
// ---------------------------------------------------------------------------
struct EosP {
    int kind;               // 0 row/col, 1 octave
    // row/col (eos_table.h:39-62)
    int nrow, ncol, num_ed;
    float ed_start, ed_step, epr, r_epr, r_step, top;
    // octave
    int kmin, kmax, lj, nlast;
    float lo_e, hi_e, r_lo, r_hi;
    const float4 *tab;
};

struct PhysP {
    int nx, ny, nz;
    int64_t n, sx, sy;      // cells, stride of i, stride of j (k stride 1)
    float dx, dy, dz, dt;
    float r_dx, r_dy, r_dz, r_dt, r_3;
    float etaos_xmin, etaos_ymin, etaos_ls, etaos_rs, lam1;
    float zmax, ztpeak, zwidth, r_zwidth;   // r_zwidth = RN(1 / zwidth), from the host
    float fp64_vmin;
    int flags;
    int lam_fast;           // 1 when the table's cs2 lie in [0, 0.5] (max(lambda,1) shortcut valid)
    uint32_t nb;            // bytes of one field plane (4 n) when 10 planes fit 32-bit offsets, else 0
};

struct StageP {
    float tau, r_tau, r_taudz, old_time, r_step;
    int step;
};

// jet energy-loss source (kernel_jet_eloss.cl:8-69): on = mask of the RK stages that add it
// (1: stage 1, 2: stage 2), 0 adds nothing.  c, sn: cos, sin of
// the direction; x0, y0, z0: coordinates of the start cell; c3 = (2 pi)^(3/2); tau0: TAU0.
struct JetP {
    int on;
    float c, sn, x0, y0, z0, c3, tau0;
};

// ---------------------------------------------------------------------------
// Elementary operations
// ---------------------------------------------------------------------------
__device__ __forceinline__ float rcp(float b) { return 1.0f / b; }     // correctly rounded
// 1/b for 2^-126 <= |b| < 2^126: the hardware reciprocal refined by one FMA Newton step
// (3 instructions against the 10 of the correctly rounded division).  On that whole range it
// equals rcp(b) for every float (cv_selftest_fastmath runs the check on the device); its
// callers pass only operands that lie in that range by construction.
__device__ __forceinline__ float rcpn(float b) {
    const float r0 = __builtin_amdgcn_rcpf(b);
    return fmaf(fmaf(-b, r0, 1.0f), r0, r0);
}
// sqrt(y) for 2^-24 <= y <= 1: the hardware reciprocal square root r, s = y r and one FMA
// Newton step (1 transcendental and 4 instructions against about 12 for the correctly
// rounded sqrtf); on that whole range it equals sqrtf(y) for every float (device check as
// for rcpn).  The bare hardware square root differs from sqrtf on about 15 % of all floats.
__device__ __forceinline__ float sqrtn(float y) {
    const float r = __builtin_amdgcn_rsqf(y);
    const float s = y * r;
    return fmaf(fmaf(-s, s, y), 0.5f * r, s);
}
__device__ __forceinline__ float rmax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float rmin(float a, float b) { return b < a ? b : a; }
// Regulation of the updated bulk pressure (the bulk Pi extension): |Pi| <= P(e)/2 and
// Pi = 0 in dilute cells (e < CV_BULK_EMIN).  pr = P(e), e = the stage's e.
__device__ __forceinline__ float bulk_regulate(float b, float pr, float e) {
    const float h = 0.5f * pr;
    b = rmin(b, h);
    b = rmax(b, -h);
    return e < CV_BULK_EMIN ? 0.0f : b;
}
// Second regulation of the bulk extension (zeta/s > 0 only): pi^{mu nu} and Pi scaled down
// together so that the largest |component| is at most the enthalpy e + P (the reference's
// bad-cell test against T^tautau alone lets |pi| reach gamma^2 (e + P) in the fast dilute
// periphery, and the bulk pressure makes that region unstable).  max_abs follows.
template <bool BULK>
__device__ __forceinline__ void enthalpy_cap(float zmax, float out[10], float &bout, float &max_abs, float e, float pr) {
    if (BULK && zmax > 0.0f) {
        const float lim = e + pr;
        if (max_abs > lim) {
            const float sc = lim * (1.0f / max_abs);
#pragma unroll
            for (int m = 0; m < 10; m++) out[m] = out[m] * sc;
            bout = bout * sc;
            max_abs = lim;
        }
    }
}
__device__ __forceinline__ float rsign(float x) {
    return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : x);
}
__device__ __forceinline__ float minmod(float x, float y) {
    float res = rmin(fabsf(x), fabsf(y));
    return res * (rsign(x) + rsign(y)) * 0.5f;
}
// generalised minmod slope at b from cells (a, b, c):
//   minmod(0.5 (c-a), minmod(THETA (c-b), THETA (b-a)))   (helper.h:213-217)
// For finite x, y, minmod(x, y) = min(|x|, |y|) (sgn x + sgn y) / 2 is the median of
// (x, y, 0): the operand of smaller magnitude when the signs agree, else zero.  The
// median is one of its operands, so med3 returns the reference's value bit for bit,
// except that a zero result may carry the other sign; c-a has the sign of c-b when c-b
// and b-a agree (the difference of two distinct floats is never zero).
__device__ __forceinline__ float slope(float a, float b, float c) {
    const float inner = __builtin_amdgcn_fmed3f(CV_THETA * (c - b), CV_THETA * (b - a), 0.0f);
    return __builtin_amdgcn_fmed3f(0.5f * (c - a), inner, 0.0f);
}
__device__ __forceinline__ float gamma3(float vx, float vy, float vz) {
    // max(1 - v^2, acu) lies in [1e-7, 1] (sqrtn's range), its square root in [3.1e-4, 1] (rcpn's)
    return rcpn(sqrtn(rmax(1.0f - vx * vx - vy * vy - vz * vz, CV_ACU)));
}

// ---------------------------------------------------------------------------
// EOS lookup
// ---------------------------------------------------------------------------
__device__ __forceinline__ float4 lerp4(float r, float4 lo, float4 hi) {
    float s = 1.0f - r;
    return make_float4(s * lo.x + r * hi.x, s * lo.y + r * hi.y,
                       s * lo.z + r * hi.z, s * lo.w + r * hi.w);
}

// reference row/col table (eos_table.h:39-62)
__device__ __forceinline__ float4 eos_rowcol(const EosP &E, float eps) {
    int row = (int)(eps * E.r_epr);
    int col = (int)((eps - E.ed_start - (float)row * E.epr) * E.r_step);
    float eps_low = E.ed_start + (float)row * E.epr + (float)col * E.ed_step;
    int c0 = col < 0 ? 0 : (col > E.ncol - 1 ? E.ncol - 1 : col);
    int r0 = row < 0 ? 0 : (row > E.nrow - 1 ? E.nrow - 1 : row);
    float4 lo = E.tab[(int64_t)r0 * E.ncol + c0];
    col = col + 1;
    if (col == E.ncol) { row += 1; col = 0; }
    int c1 = col < 0 ? 0 : (col > E.ncol - 1 ? E.ncol - 1 : col);
    int r1 = row < 0 ? 0 : (row > E.nrow - 1 ? E.nrow - 1 : row);
    float4 hi = E.tab[(int64_t)r1 * E.ncol + c1];
    float r = (eps - eps_low) * E.r_step;
    return lerp4(r, lo, hi);
}

// octave table: node and weight from the binary exponent and mantissa of eps
__device__ __forceinline__ float4 eos_octave(const EosP &E, const float4 *tab, float eps) {
    if (!(eps >= E.lo_e)) {
        float r = eps > 0.0f ? eps * E.r_lo : 0.0f;
        float4 t = tab[0];
        return make_float4(t.x, r * t.y, r * t.z, r * t.w);
    }
    if (eps >= E.hi_e) {
        float r = eps * E.r_hi;
        float4 t = tab[E.nlast];
        return make_float4(t.x, r * t.y, r * t.z, r * t.w);
    }
    int ex;
    float fr = frexpf(eps, &ex);
    float f = 2.0f * fr - 1.0f;
    float tt = f * (float)(1 << E.lj);
    int j = (int)tt;
    float r = tt - (float)j;
    int n = ((ex - 1 - E.kmin) << E.lj) + j;
    return lerp4(r, tab[n], tab[n + 1]);
}

// A lookup context: the table pointer may be an LDS copy of E.tab.
// K = 1: octave table, K = 0: row/col table, K = -1: chosen at run time from E.kind
// (kernels specialised on the table kind carry only that kind's parameters)
template <int K> struct EosT {
    const EosP &E;
    const float4 *tab;
    __device__ EosT(const EosP &e, const float4 *t) : E(e), tab(t) {}
    __device__ __forceinline__ bool octave() const { return K == 1 || (K < 0 && E.kind == 1); }
    __device__ __forceinline__ float4 raw(float eps) const {
        return octave() ? eos_octave(E, tab, eps) : eos_rowcol(E, eps);
    }
    __device__ __forceinline__ bool in_table(float eps) const {
        return octave() || eps < E.top;
    }
    // P, S, T, CS2 with the analytic fit above a row/col table (eos_table.h:65-100)
    __device__ __forceinline__ float P(float eps) const {
        if (in_table(eps)) return raw(eps).y;
        return 0.3327f * eps - 0.3223f * powf(eps, 0.4585f) - 0.003906f * eps * expf(-0.05697f * eps);
    }
    __device__ __forceinline__ float S(float eps) const {
        if (in_table(eps)) return raw(eps).w;
        return 18.202f * eps - 63.0218f - 4.85479f * expf(-2.72407E-11f * powf(eps, 4.54886f))
               + 65.1272f * powf(eps, -0.128012f) * expf(-0.00369624f * powf(eps, 1.18735f));
    }
    __device__ __forceinline__ float T(float eps) const {
        if (in_table(eps)) return raw(eps).z;
        return (eps + P(eps)) * rcp(S(eps));
    }
    __device__ __forceinline__ float CS2(float eps) const {
        if (in_table(eps)) return raw(eps).x;
        return 0.333333333f;
    }
};
using Eos = EosT<-1>;

// Octave table whose pressure column (P at the nodes) is held in LDS (pt, nlast + 1 floats):
// P() reads the column, the other functions the whole table through the caches.  P is the
// .y component of eos_octave's lerp bit for bit (same node, weight and operations), so the
// results are those of EosT<1>; only the latency of the pressure lookups changes.
struct EosOctP {
    const EosP &E;
    const float4 *tab;
    const float *pt;
    __device__ EosOctP(const EosP &e, const float4 *t, const float *p_) : E(e), tab(t), pt(p_) {}
    __device__ __forceinline__ bool octave() const { return true; }
    __device__ __forceinline__ float4 raw(float eps) const { return eos_octave(E, tab, eps); }
    __device__ __forceinline__ bool in_table(float) const { return true; }
    __device__ __forceinline__ float P(float eps) const {
        if (!(eps >= E.lo_e)) {
            float r = eps > 0.0f ? eps * E.r_lo : 0.0f;
            return r * pt[0];
        }
        if (eps >= E.hi_e) return (eps * E.r_hi) * pt[E.nlast];
        int ex;
        float fr = frexpf(eps, &ex);
        float f = 2.0f * fr - 1.0f;
        float tt = f * (float)(1 << E.lj);
        int j = (int)tt;
        float r = tt - (float)j;
        int n = ((ex - 1 - E.kmin) << E.lj) + j;
        float s = 1.0f - r;
        return s * pt[n] + r * pt[n + 1];
    }
    __device__ __forceinline__ float S(float eps) const { return raw(eps).w; }
    __device__ __forceinline__ float T(float eps) const { return raw(eps).z; }
    __device__ __forceinline__ float CS2(float eps) const { return raw(eps).x; }
};

// ---------------------------------------------------------------------------
// Hydro helpers (helper.h)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float etaos(const PhysP &p, float t) {
    return (t < p.etaos_xmin) ? (p.etaos_ls * (t - p.etaos_xmin) + p.etaos_ymin)
                              : (p.etaos_rs * (t - p.etaos_xmin) + p.etaos_ymin);
}
__device__ __forceinline__ float zetaos(const PhysP &p, float t) {
    float d = (t - p.ztpeak) * p.r_zwidth;
    return p.zmax * rcp(1.0f + d * d);
}

// T^{tau mu} (helper.h:172-177); g = gamma of ev
__device__ __forceinline__ float4 t0m(float4 ev, float pr) {
    float u0 = gamma3(ev.y, ev.z, ev.w);
    float a = (ev.x + pr) * u0;
    return make_float4(a * (1.0f * u0) - 1.0f * pr, a * (ev.y * u0) - 0.0f * pr,
                       a * (ev.z * u0) - 0.0f * pr, a * (ev.w * u0) - 0.0f * pr);
}

// maxPropagationSpeed (helper.h:94-105) = max(lambda_KT, 1).  Take X = ut^2 >= 1 as an
// input of its own (gamma3 never returns less than 1, whatever v is), a = 1 - cs2, c = cs:
//   1 - lambda_KT = X a (1 - |vk|)^2 / (X a (1 - |vk|) + cs2 + c sqrt(X (1 - vk^2) a + cs2)),
// which is >= 0 for every X >= 0, |vk| <= 1 and 0 <= cs2 <= 1.  With 0 <= cs2 <= 0.5 and
// vk^2 <= 0.98 it is at least 5.0e-5 (839 FP32 ulps of 1, the minimum at X = 1,
// cs2 = 0.5), while the FP32 evaluation of the formula (every operation rounded once) is
// off by a few ulps whatever X is (at most 6.3 ulps over 1.2e8 samples with X up to 1e7).
// So the maximum is exactly 1.0f there, and only the face velocity component is tested.
// lam_fast: every cs2 the lookup can return lies in [0, 0.5] (checked on the table).
template <class EOS>
__device__ __forceinline__ float max_prop_speed(const PhysP &p, const EOS &eos, float4 edv, float vk) {
    if (p.lam_fast && vk * vk <= 0.98f) return 1.0f;
    float ut = gamma3(edv.y, edv.z, edv.w);
    float uk = ut * vk;
    float ut2 = ut * ut;
    float uk2 = uk * uk;
    float cs2 = eos.CS2(edv.x);
    float lam = (fabsf(ut * uk * (1.0f - cs2)) + sqrtf((ut2 - uk2 - (ut2 - uk2 - 1.0f) * cs2) * cs2))
                * rcp(ut2 - (ut2 - 1.0f) * cs2);
    return rmax(lam, 1.0f);
}

// rootFinding_newton (helper.h:128-167) with P_eff = P + bulk
// bulk: Pi added to P(e); clamp0: effective pressure P + Pi held >= 0 (bulk runs only:
// a negative effective pressure would allow |v| > 1 in v = M / (T00 + P + Pi)).
// v0: start of the iteration (the reference starts at the middle of [0, 1]; a learned
// seed starts elsewhere, the bracket and the stopping rule stay the same).
// RN: the Newton step divides with rcpn.  Valid for the stages' own states: T00 >= 1e-7 / tau
// (the acu floor), M <= 0.999 T00 and P + Pi >= 0, so df >= (2/3) T00 is in rcpn's range;
// states given from outside (cv_root_eval) take the correctly rounded division.
template <bool RN = true, class EOS>
__device__ __forceinline__ float root_newton_from(const EOS &eos, float T00, float M, float bulk, bool clamp0,
                                                  float v0, int *iters) {
    float vl = 0.0f, vh = 1.0f;
    const float dpe = 1.0f / 3.0f;
    float v = v0;
    float ed = T00 - M * v;
    float pr = eos.P(ed) + bulk;
    if (clamp0) pr = rmax(pr, 0.0f);
    float f = (T00 + pr) * v - M;
    float df = (T00 + pr) - M * v * dpe;
    float dvold = vh - vl;
    float dv = dvold;
    int i = 0;
    while (true) {
        if ((f + df * (vh - v)) * (f + df * (vl - v)) > 0.0f || fabsf(2.f * f) > fabsf(dvold * df)) {
            dvold = dv;
            dv = 0.5f * (vh - vl);
            v = vl + dv;
        } else {
            dvold = dv;
            dv = f * (RN ? rcpn(df) : rcp(df));
            v -= dv;
        }
        i++;
        if (fabsf(dv) < 0.00001f || i > 100) break;
        ed = T00 - M * v;
        pr = eos.P(ed) + bulk;
        if (clamp0) pr = rmax(pr, 0.0f);
        f = (T00 + pr) * v - M;
        df = (T00 + pr) - M * v * dpe;
        if (f > 0.0f) vh = v; else vl = v;
    }
    if (iters) *iters = i;
    return T00 - M * v;
}
template <class EOS>
__device__ __forceinline__ float root_newton(const EOS &eos, float T00, float M, float bulk, bool clamp0 = false) {
    return root_newton_from(eos, T00, M, bulk, clamp0, 0.5f * (0.0f + 1.0f), nullptr);
}

// FP64 twin of the root solve (cells with M close to T00); table values stay FP32.
__device__ __forceinline__ double eos_P_d(const EosP &E, const float4 *tab, double eps) {
    if (E.kind == 1) {
        if (!(eps >= (double)E.lo_e)) return (eps > 0.0 ? eps * (double)E.r_lo : 0.0) * (double)tab[0].y;
        if (eps >= (double)E.hi_e) return eps * (double)E.r_hi * (double)tab[E.nlast].y;
        int ex;
        double fr = frexp(eps, &ex);
        double f = 2.0 * fr - 1.0;
        double tt = f * (double)(1 << E.lj);
        int j = (int)tt;
        double r = tt - j;
        int n = ((ex - 1 - E.kmin) << E.lj) + j;
        return (1.0 - r) * (double)tab[n].y + r * (double)tab[n + 1].y;
    }
    double step = E.ed_step, start = E.ed_start;
    double epr = step * E.ncol;
    if (eps < (double)E.ed_start + (double)E.num_ed * (double)E.ed_step) {
        int row = (int)(eps * (1.0 / epr));
        int col = (int)((eps - start - row * epr) * (1.0 / step));
        double elow = start + row * epr + col * step;
        int c0 = col < 0 ? 0 : (col > E.ncol - 1 ? E.ncol - 1 : col);
        int r0 = row < 0 ? 0 : (row > E.nrow - 1 ? E.nrow - 1 : row);
        double lo = tab[(int64_t)r0 * E.ncol + c0].y;
        col += 1;
        if (col == E.ncol) { row += 1; col = 0; }
        int c1 = col < 0 ? 0 : (col > E.ncol - 1 ? E.ncol - 1 : col);
        int r1 = row < 0 ? 0 : (row > E.nrow - 1 ? E.nrow - 1 : row);
        double hi = tab[(int64_t)r1 * E.ncol + c1].y;
        double r = (eps - elow) * (1.0 / step);
        return (1.0 - r) * lo + r * hi;
    }
    return 0.3327 * eps - 0.3223 * pow(eps, 0.4585) - 0.003906 * eps * exp(-0.05697 * eps);
}

// The kernels that finish deferred cells (k_fp64_fix, k_batch_fp64_fix) and the split root
// kernels (k_visc_root, k_batch_visc_root) inline it; the per-cell stage kernels that carry
// the branch call root_newton_d, out of line: few cells take it there, and inlined its
// double registers would spill.
__device__ __forceinline__ double root_newton_d_inl(const EosP &E, const float4 *tab, double T00, double M,
                                                    double bulk, bool clamp0 = false) {
    double vl = 0.0, vh = 1.0, dpe = 1.0 / 3.0;
    double v = 0.5;
    double ed = T00 - M * v;
    double pr = eos_P_d(E, tab, ed) + bulk;
    if (clamp0) pr = pr > 0.0 ? pr : 0.0;
    double f = (T00 + pr) * v - M;
    double df = (T00 + pr) - M * v * dpe;
    double dvold = vh - vl, dv = dvold;
    int i = 0;
    while (true) {
        if ((f + df * (vh - v)) * (f + df * (vl - v)) > 0.0 || fabs(2.0 * f) > fabs(dvold * df)) {
            dvold = dv; dv = 0.5 * (vh - vl); v = vl + dv;
        } else {
            dvold = dv; dv = f * (1.0 / df); v -= dv;
        }
        i++;
        if (fabs(dv) < 0.00001 || i > 100) break;
        ed = T00 - M * v;
        pr = eos_P_d(E, tab, ed) + bulk;
        if (clamp0) pr = pr > 0.0 ? pr : 0.0;
        f = (T00 + pr) * v - M;
        df = (T00 + pr) - M * v * dpe;
        if (f > 0.0) vh = v; else vl = v;
    }
    return T00 - M * v;
}
__device__ __noinline__ double root_newton_d(const EosP &E, const float4 *tab, double T00, double M, double bulk,
                                             bool clamp0 = false) {
    return root_newton_d_inl(E, tab, T00, M, bulk, clamp0);
}

}  // namespace cv
