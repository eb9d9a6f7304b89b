/*
 * clvisc_oracle.c -- CPU oracle for the CLVisc (3+1)D KT + Israel-Stewart hydro step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the arithmetic of
 * the reference OpenCL kernels and the RK2 sequencing of the reference Python
 * drivers, so that the HIP product path can be checked against it.  Only the
 * test suite, the smoke check and the CPU-baseline leg of bench.py may load it;
 * the product path never calls, links or falls back to it.
 *
 * Reference files restated (paths relative to the reference repository):
 *   pyvisc/kernel/helper.h        gamma, umu4, t0m, minmod, maxPropagationSpeed,
 *                                 rootFinding_newton, kt1d, kt1d_real, etaos
 *   pyvisc/kernel/eos_table.h     eos(), P, S, T, CS2 (uniform row/col table)
 *   pyvisc/kernel/kernel_ideal.cl kt_src_christoffel, kt_src_along{x,y,z}, update_ev
 *   pyvisc/kernel/kernel_visc.cl  kt_src_christoffel, kt1d_visc, kt_src_along*, update_ev
 *   pyvisc/kernel/kernel_IS.cl    visc_initialize, visc_src_christoffel,
 *                                 visc_src_along*, update_pimn, get_udiff
 *   pyvisc/kernel/kernel_reduction.cl  reduction_stage1 (max of e)
 *   pyvisc/ideal.py:211-255, 317-349   CLIdeal.stepUpdate / update_time / evolve
 *   pyvisc/visc.py:143-149, 232-314, 352-357, 468-560  CLVisc stages and loop
 *
 * Each function keeps the reference's operation order (left-to-right, no FMA:
 * build with -ffp-contract=off) so that FP32 rounding follows the reference as
 * closely as a different compiler allows.  Compiled twice: ORC_DOUBLE=0 gives the
 * FP32 oracle, ORC_DOUBLE=1 the FP64 one (same algorithm in double; acu=1e-14
 * as in pyvisc/kernel/real_type.h:16).
 *
 * Extensions beyond the reference (each off unless its flag is set, and each
 * reduces exactly to the reference path when off):
 *   eos_kind = 1      octave table (uniform nodes inside each binary octave of e,
 *                     indexed from the float's exponent and mantissa) instead of
 *                     the row/col table
 *   bulk_on           bulk viscous pressure Pi with zeta/s(T) (absent in the reference)
 *   fp64_root         conserved-to-primitive solve in double where M > vmin*T00
 *   plant             tests only: a mistake planted in the eta branches (ORC_PLANT_*),
 *                     which the eta-direction pins must detect; 0 everywhere else
 *
 * Layouts are the reference's AoS: ev[N][4] = (e, vx, vy, veta) with veta = tau*deta/dtau,
 * pi[N][10] in idx() order (00,01,02,03,11,12,13,22,23,33) with eta components
 * tilde-scaled, cell index I = (i*NY + j)*NZ + k.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORC_DOUBLE
#define ORC_DOUBLE 0
#endif

#if ORC_DOUBLE
typedef double real;
#define ACU 1.0e-14
#else
typedef float real;
#define ACU 1.0e-7f
#endif
#define RC(x) ((real)(x))

/* Arithmetic contract shared with the HIP path.  The reference's OpenCL C allows
 * single-precision division within 2.5 ulp (no -cl-fp32-correctly-rounded-divide-sqrt
 * in pyvisc/ideal.py:146-171); every quotient a/b is evaluated here as a * RN(1/b),
 * which lies within 1.5 ulp, with RN(1/b) the correctly rounded reciprocal.  sqrt is
 * correctly rounded, nothing is contracted into FMA.  The HIP kernels evaluate the
 * same operations in the same order, so their FP32 results match this oracle bit for
 * bit on the division-free EOS table (rcp of a run constant is computed once). */
static real rcp(real b) { return RC(1.0f) / b; }
#define DIV(a, b) ((a) * rcp(b))

/* THETA and hbarc are float literals in the reference (helper.h:7-8). */
#define THETA RC(1.1f)
#if ORC_DOUBLE
#define R_FMIN fmin
#define R_COPYSIGN copysign
#else
#define R_FMIN fminf
#define R_COPYSIGN copysignf
#endif
#define HBARC RC(0.19733f)

/* ------------------------------------------------------------------------- */
/* Parameters shared with the Python wrapper (oracle/oracle.py).              */
/* ------------------------------------------------------------------------- */
typedef struct orc_params {
    int32_t nx, ny, nz;
    int32_t riemann_test;      /* -D RIEMANN_TEST */
    int32_t gubser_visc_test;  /* -D GUBSER_VISC_TEST */
    int32_t eos_kind;          /* 0: reference row/col table, 1: octave table */
    int32_t eos_nrow, eos_ncol, eos_num_ed;   /* kind 0 */
    int32_t eos_kmin, eos_kmax, eos_log2j;    /* kind 1 */
    float dt, dx, dy, dz, tau0;
    float etaos_xmin, etaos_ymin, etaos_left_slop, etaos_right_slop, lam1;
    float eos_ed_start, eos_ed_step;          /* kind 0 */
    const float *eos_table;                   /* RGBA = (cs2, P, T, s) */
    int32_t bulk_on;
    float zetaos_max, zetaos_tpeak, zetaos_width;
    int32_t fp64_root;
    float fp64_root_vmin;
    int32_t nthreads;          /* OpenMP threads, 0 = runtime default */
    double tau0_cfg, dt_cfg;   /* TAU0, DT as the Python config holds them (double) */
    int32_t couplings;         /* 1: -D PIMUNU_OMEGA_COUPLING, 2: -D OMEGA_OMEGA_COUPLING */
    int32_t jet_on;            /* RK stages given jet_eloss_src (1: stage 1, 2: stage 2) */
    float jet_direction;       /* jet_eloss_src['direction'] (cfg.real) */
    int64_t jet_start;         /* jet_eloss_src['start_pos_index'] */
    int32_t plant;             /* tests only: mistakes planted to show that the eta pins fail
                                  (ORC_PLANT_*); 0 in every other use */
} orc_params;

/* Planted mistakes (tests/test_eta_oracle.py shows that each fails the eta-direction pins):
 * the eta sweeps divide by DZ instead of tau*DZ; udz without its Christoffel term. */
#define ORC_PLANT_ETA_DZ 1
#define ORC_PLANT_UDZ_NO_CHRISTOFFEL 2

static int64_t ncells(const orc_params *p) { return (int64_t)p->nx * p->ny * p->nz; }

/* ------------------------------------------------------------------------- */
/* EOS lookup                                                                 */
/* ------------------------------------------------------------------------- */

/* Reference row/col table lookup with nearest-texel clamp-to-edge reads and a
 * manual lerp.  Follows pyvisc/kernel/eos_table.h:39-62. */
static void eos_rowcol(const orc_params *p, real eps, real out[4]) {
    const real step = RC(p->eos_ed_step);
    const real start = RC(p->eos_ed_start);
    const int ncol = p->eos_ncol, nrow = p->eos_nrow;
    real ed_per_row = step * ncol;
    int row = (int)DIV(eps, ed_per_row);
    int col = (int)DIV(eps - start - row * ed_per_row, step);
    real eps_low = start + row * ed_per_row + col * step;
    int c0 = col < 0 ? 0 : (col > ncol - 1 ? ncol - 1 : col);
    int r0 = row < 0 ? 0 : (row > nrow - 1 ? nrow - 1 : row);
    const float *lo = p->eos_table + 4 * ((int64_t)r0 * ncol + c0);
    col = col + 1;
    if (col == ncol) { row += 1; col = 0; }
    int c1 = col < 0 ? 0 : (col > ncol - 1 ? ncol - 1 : col);
    int r1 = row < 0 ? 0 : (row > nrow - 1 ? nrow - 1 : row);
    const float *hi = p->eos_table + 4 * ((int64_t)r1 * ncol + c1);
    real r = DIV(eps - eps_low, step);
    for (int c = 0; c < 4; c++) out[c] = (RC(1.0f) - r) * RC(lo[c]) + r * RC(hi[c]);
}

/* Octave table (extension): nodes at e = 2^k (1 + j/J) for k in [kmin, kmax),
 * j in [0, J), J = 2^log2j, plus a last node at 2^kmax; node n = (k-kmin) J + j.
 * The node and the interpolation weight come straight from the binary exponent
 * and mantissa of e (exact, no log and no division); values are interpolated
 * linearly in e.  Below 2^kmin P, T, s fall linearly to zero at e = 0 (cs2 held);
 * at or above 2^kmax they scale linearly with e from the last node. */
static void eos_oct(const orc_params *p, real eps, real out[4]) {
    const float *t = p->eos_table;
    const int kmin = p->eos_kmin, kmax = p->eos_kmax, lj = p->eos_log2j;
    const int nlast = (kmax - kmin) << lj;
    if (!(eps >= (real)ldexp(1.0, kmin))) {
        real r = eps > RC(0.0f) ? eps * (real)ldexp(1.0, -kmin) : RC(0.0f);
        out[0] = RC(t[0]);
        for (int c = 1; c < 4; c++) out[c] = r * RC(t[c]);
        return;
    }
    if (eps >= (real)ldexp(1.0, kmax)) {
        real r = eps * (real)ldexp(1.0, -kmax);
        out[0] = RC(t[4 * nlast]);
        for (int c = 1; c < 4; c++) out[c] = r * RC(t[4 * nlast + c]);
        return;
    }
    int ex;
#if ORC_DOUBLE
    real fr = frexp(eps, &ex);
#else
    real fr = frexpf(eps, &ex);
#endif
    /* eps = 2^k (1 + f), f in [0, 1) */
    int k = ex - 1;
    real f = RC(2.0f) * fr - RC(1.0f);
    real tt = f * (real)(1 << lj);
    int j = (int)tt;
    real r = tt - (real)j;
    int n = ((k - kmin) << lj) + j;
    const float *lo = t + 4 * n, *hi = lo + 4;
    for (int c = 0; c < 4; c++) out[c] = (RC(1.0f) - r) * RC(lo[c]) + r * RC(hi[c]);
}

static void eos4(const orc_params *p, real eps, real out[4]) {
    if (p->eos_kind == 1) eos_oct(p, eps, out);
    else eos_rowcol(p, eps, out);
}

static int in_table(const orc_params *p, real eps) {
    if (p->eos_kind == 1) return 1;
    return eps < RC(p->eos_ed_start) + p->eos_num_ed * RC(p->eos_ed_step);
}

/* P, S, T, CS2 with the analytic fit above the table (eos_table.h:65-100). */
static real eos_P(const orc_params *p, real eps) {
    if (in_table(p, eps)) { real o[4]; eos4(p, eps, o); return o[1]; }
    return RC(0.3327f) * eps - RC(0.3223f) * (real)pow((double)eps, (double)RC(0.4585f))
           - RC(0.003906f) * eps * (real)exp((double)(RC(-0.05697f) * eps));
}
static real eos_S(const orc_params *p, real eps) {
    if (in_table(p, eps)) { real o[4]; eos4(p, eps, o); return o[3]; }
    return RC(18.202f) * eps - RC(63.0218f)
           - RC(4.85479f) * (real)exp((double)(RC(-2.72407E-11f) * (real)pow((double)eps, (double)RC(4.54886f))))
           + RC(65.1272f) * (real)pow((double)eps, (double)(-RC(0.128012f)))
                 * (real)exp((double)(RC(-0.00369624f) * (real)pow((double)eps, (double)RC(1.18735f))));
}
static real eos_T(const orc_params *p, real eps) {
    if (in_table(p, eps)) { real o[4]; eos4(p, eps, o); return o[2]; }
    return DIV(eps + eos_P(p, eps), eos_S(p, eps));
}
static real eos_CS2(const orc_params *p, real eps) {
    if (in_table(p, eps)) { real o[4]; eos4(p, eps, o); return o[0]; }
    return RC(0.333333333f);
}

/* ------------------------------------------------------------------------- */
/* Pointwise helpers (helper.h)                                               */
/* ------------------------------------------------------------------------- */
static real r_max(real a, real b) { return a > b ? a : b; }
static real r_min(real a, real b) { return b < a ? b : a; }   /* OpenCL min(): y if y < x */
static real r_sign(real x) {
    if (x > RC(0.0f)) return RC(1.0f);
    if (x < RC(0.0f)) return RC(-1.0f);
    return x;  /* keeps +-0 */
}

/* etaos(T) piecewise linear, helper.h:23-26 */
static real etaos(const orc_params *p, real t) {
    return (t < RC(p->etaos_xmin)) ? (RC(p->etaos_left_slop) * (t - RC(p->etaos_xmin)) + RC(p->etaos_ymin))
                                   : (RC(p->etaos_right_slop) * (t - RC(p->etaos_xmin)) + RC(p->etaos_ymin));
}

/* zeta/s(T) for the bulk extension: Cauchy peak at zetaos_tpeak. */
static real zetaos(const orc_params *p, real t) {
    real d = DIV(t - RC(p->zetaos_tpeak), RC(p->zetaos_width));
    return DIV(RC(p->zetaos_max), RC(1.0f) + d * d);
}

static real gamma3(real vx, real vy, real vz) {
    return rcp((real)sqrt((double)r_max(RC(1.0f) - vx * vx - vy * vy - vz * vz, ACU)));
}
static real gamma_ev(const real *ev) { return gamma3(ev[1], ev[2], ev[3]); }

static void umu4(const real *ev, real u[4]) {
    real g = gamma_ev(ev);
    u[0] = g * RC(1.0f); u[1] = g * ev[1]; u[2] = g * ev[2]; u[3] = g * ev[3];
}

static real minmod(real x, real y) {
    real res = r_min((real)fabs((double)x), (real)fabs((double)y));
    return res * (r_sign(x) + r_sign(y)) * RC(0.5f);
}

/* generalised minmod slope at b from (a, b, c) = cells (i-1, i, i+1), as written in
 * kt1d / kt1d_visc / kt1d_real (helper.h:213-217, kernel_visc.cl:80-84):
 *   minmod(0.5 (c-a), minmod(THETA (c-b), THETA (b-a))).
 * Evaluated in closed form: 0 when c-b and b-a carry different sign bits, otherwise
 * copysign(min(0.5|c-a|, THETA|c-b|, THETA|b-a|), c-b) (a zero difference makes the
 * min 0).  Same value bit for bit for finite inputs (every product with 0.5 or a sign
 * is exact); only the sign of an exact zero may differ. */
static real slope(real a, real b, real c) {
    real d2 = c - b, d3 = b - a;
    real m = R_FMIN(RC(0.5f) * (real)fabs((double)(c - a)),
                    R_FMIN(THETA * (real)fabs((double)d2), THETA * (real)fabs((double)d3)));
    if (signbit(d2) != signbit(d3)) return RC(0.0f);
    return R_COPYSIGN(m, d2);
}

/* T^{tau mu} = (e+P) u^tau u^mu - P g^{tau mu}  (helper.h:172-177) */
static void t0m(const real *ev, real pr, real out[4]) {
    real u0 = gamma3(ev[1], ev[2], ev[3]);
    real ed = ev[0];
    real u4[4] = {RC(1.0f) * u0, ev[1] * u0, ev[2] * u0, ev[3] * u0};
    real a = (ed + pr) * u0;
    out[0] = a * u4[0] - RC(1.0f) * pr;
    out[1] = a * u4[1] - RC(0.0f) * pr;
    out[2] = a * u4[2] - RC(0.0f) * pr;
    out[3] = a * u4[3] - RC(0.0f) * pr;
}

/* maxPropagationSpeed (helper.h:94-105): returns max(lambda, 1). */
static real max_prop_speed(const orc_params *p, const real *edv, real vk) {
    real ut = gamma3(edv[1], edv[2], edv[3]);
    real uk = ut * vk;
    real ut2 = ut * ut;
    real uk2 = uk * uk;
    real cs2 = eos_CS2(p, edv[0]);
    real lam = DIV((real)fabs((double)(ut * uk * (RC(1.0f) - cs2)))
                   + (real)sqrt((double)((ut2 - uk2 - (ut2 - uk2 - RC(1.0f)) * cs2) * cs2)),
                   ut2 - (ut2 - RC(1.0f)) * cs2);
    return r_max(lam, RC(1.0f));
}

/* rootFinding_newton (helper.h:128-167) with P_eff(e) = P(e) + bulk; clamp0 (bulk
 * runs) holds P_eff >= 0 so that v = M / (T00 + P_eff) stays below 1. */
static real root_newton_v0(const orc_params *p, real T00, real M, real bulk, int clamp0, real v0, int *iters);
static real root_newton(const orc_params *p, real T00, real M, real bulk, int clamp0, int *iters) {
    return root_newton_v0(p, T00, M, bulk, clamp0, RC(0.5f) * (RC(0.0f) + RC(1.0f)), iters);
}
/* the same iteration from a given start v0 (the learned seed; no reference counterpart) */
static real root_newton_v0(const orc_params *p, real T00, real M, real bulk, int clamp0, real v0, int *iters) {
    real vl = RC(0.0f), vh = RC(1.0f);
    real dpe = RC(1.0f) / RC(3.0f);
    real v = v0;
    real ed = T00 - M * v;
    real pr = eos_P(p, ed) + bulk;
    if (clamp0) pr = r_max(pr, RC(0.0f));
    real f = (T00 + pr) * v - M;
    real df = (T00 + pr) - M * v * dpe;
    real dvold = vh - vl;
    real dv = dvold;
    int i = 0;
    while (1) {
        if ((f + df * (vh - v)) * (f + df * (vl - v)) > RC(0.0f) ||
            (real)fabs((double)(RC(2.f) * f)) > (real)fabs((double)(dvold * df))) {
            dvold = dv;
            dv = RC(0.5f) * (vh - vl);
            v = vl + dv;
        } else {
            dvold = dv;
            dv = DIV(f, df);
            v -= dv;
        }
        i++;
        if ((real)fabs((double)dv) < RC(0.00001f) || i > 100) break;
        ed = T00 - M * v;
        pr = eos_P(p, ed) + bulk;
        if (clamp0) pr = r_max(pr, RC(0.0f));
        f = (T00 + pr) * v - M;
        df = (T00 + pr) - M * v * dpe;
        if (f > RC(0.0f)) vh = v; else vl = v;
    }
    if (iters) *iters = i;
    return T00 - M * v;
}

#if !ORC_DOUBLE
/* FP64 variant of the same iteration for cells with M close to T00 (extension). */
static double eos_P_d(const orc_params *p, double eps) {
    /* same lookups as eos_P with double arithmetic on the FP32 table values */
    if (p->eos_kind == 1) {
        const float *t = p->eos_table;
        const int kmin = p->eos_kmin, kmax = p->eos_kmax, lj = p->eos_log2j;
        const int nlast = (kmax - kmin) << lj;
        if (!(eps >= ldexp(1.0, kmin))) return (eps > 0.0 ? eps * ldexp(1.0, -kmin) : 0.0) * t[1];
        if (eps >= ldexp(1.0, kmax)) return eps * ldexp(1.0, -kmax) * t[4 * nlast + 1];
        int ex;
        double fr = frexp(eps, &ex);
        double f = 2.0 * fr - 1.0;
        double tt = f * (double)(1 << lj);
        int j = (int)tt;
        double r = tt - j;
        int n = ((ex - 1 - kmin) << lj) + j;
        return (1.0 - r) * t[4 * n + 1] + r * t[4 * n + 5];
    }
    if (eps < (double)p->eos_ed_start + p->eos_num_ed * (double)p->eos_ed_step) {
        double step = p->eos_ed_step, start = p->eos_ed_start;
        int ncol = p->eos_ncol, nrow = p->eos_nrow;
        double epr = step * ncol;
        int row = (int)(eps * (1.0 / epr));
        int col = (int)((eps - start - row * epr) * (1.0 / step));
        double elow = start + row * epr + col * step;
        int c0 = col < 0 ? 0 : (col > ncol - 1 ? ncol - 1 : col);
        int r0 = row < 0 ? 0 : (row > nrow - 1 ? nrow - 1 : row);
        double lo = p->eos_table[4 * ((int64_t)r0 * ncol + c0) + 1];
        col += 1;
        if (col == ncol) { row += 1; col = 0; }
        int c1 = col < 0 ? 0 : (col > ncol - 1 ? ncol - 1 : col);
        int r1 = row < 0 ? 0 : (row > nrow - 1 ? nrow - 1 : row);
        double hi = p->eos_table[4 * ((int64_t)r1 * ncol + c1) + 1];
        double r = (eps - elow) * (1.0 / step);
        return (1.0 - r) * lo + r * hi;
    }
    return 0.3327 * eps - 0.3223 * pow(eps, 0.4585) - 0.003906 * eps * exp(-0.05697 * eps);
}

static double root_newton_d(const orc_params *p, double T00, double M, double bulk, int clamp0, int *iters) {
    double vl = 0.0, vh = 1.0, dpe = 1.0 / 3.0;
    double v = 0.5;
    double ed = T00 - M * v;
    double pr = eos_P_d(p, ed) + bulk;
    if (clamp0) pr = pr > 0.0 ? pr : 0.0;
    double f = (T00 + pr) * v - M;
    double df = (T00 + pr) - M * v * dpe;
    double dvold = vh - vl, dv = dvold;
    int i = 0;
    while (1) {
        if ((f + df * (vh - v)) * (f + df * (vl - v)) > 0.0 || fabs(2.0 * f) > fabs(dvold * df)) {
            dvold = dv; dv = 0.5 * (vh - vl); v = vl + dv;
        } else {
            dvold = dv; dv = f * (1.0 / df); v -= dv;
        }
        i++;
        if (fabs(dv) < 0.00001 || i > 100) break;
        ed = T00 - M * v;
        pr = eos_P_d(p, ed) + bulk;
        if (clamp0) pr = pr > 0.0 ? pr : 0.0;
        f = (T00 + pr) * v - M;
        df = (T00 + pr) - M * v * dpe;
        if (f > 0.0) vh = v; else vl = v;
    }
    if (iters) *iters = i;
    return T00 - M * v;
}
#endif

/* solve dispatch: FP64 for fast cells when the extension is on */
static real solve_ed(const orc_params *p, real T00, real M, real bulk, int clamp0, int *iters) {
#if !ORC_DOUBLE
    if (p->fp64_root && M > RC(p->fp64_root_vmin) * T00)
        return (real)root_newton_d(p, (double)T00, (double)M, (double)bulk, clamp0, iters);
#endif
    return root_newton(p, T00, M, bulk, clamp0, iters);
}

/* ------------------------------------------------------------------------- */
/* Small exported helpers for unit tests                                      */
/* ------------------------------------------------------------------------- */
void orc_slope(const real *a, const real *b, const real *c, int64_t n, real *out) {
    for (int64_t i = 0; i < n; i++) out[i] = slope(a[i], b[i], c[i]);
}

void orc_eos(const orc_params *p, const real *eps, int64_t n, real *out /* n x 4 */) {
    for (int64_t i = 0; i < n; i++) eos4(p, eps[i], out + 4 * i);
}
void orc_eos_funcs(const orc_params *p, const real *eps, int64_t n, real *out /* n x 4: P,S,T,CS2 */) {
    for (int64_t i = 0; i < n; i++) {
        out[4 * i + 0] = eos_P(p, eps[i]);
        out[4 * i + 1] = eos_S(p, eps[i]);
        out[4 * i + 2] = eos_T(p, eps[i]);
        out[4 * i + 3] = eos_CS2(p, eps[i]);
    }
}
void orc_minmod4(const real *a, const real *b, real *out) {
    for (int c = 0; c < 4; c++) out[c] = minmod(a[c], b[c]);
}
void orc_root(const orc_params *p, const real *T00, const real *M, int64_t n, real *ed, int32_t *iters) {
    for (int64_t i = 0; i < n; i++) {
        int it = 0;
        ed[i] = solve_ed(p, T00[i], M[i], RC(0.0f), 0, &it);
        if (iters) iters[i] = it;
    }
}
void orc_root_from(const orc_params *p, const real *T00, const real *M, const real *v0, int64_t n, real *ed,
                   int32_t *iters) {
    for (int64_t i = 0; i < n; i++) {
        int it = 0;
        real s = v0[i] > RC(0.0f) ? v0[i] : RC(0.0f);
        s = s < RC(0.9999f) ? s : RC(0.9999f);
        ed[i] = root_newton_v0(p, T00[i], M[i], RC(0.0f), 0, s, &it);
        if (iters) iters[i] = it;
    }
}
void orc_max_prop_speed(const orc_params *p, const real *ev, const real *vk, int64_t n, real *out) {
    for (int64_t i = 0; i < n; i++) out[i] = max_prop_speed(p, ev + 4 * i, vk[i]);
}

/* reduction_stage1 + host max (kernel_reduction.cl:7-39, ideal.py:257-262) */
real orc_max_ed(const orc_params *p, const real *ev) {
    int64_t n = ncells(p);
    real m = RC(-3.402823466e+38f);
    for (int64_t i = 0; i < n; i++) if (ev[4 * i] > m) m = ev[4 * i];
    return m;
}

/* pizz_o_ep (kernel_checkpi.cl:5-27, visc.py:200-208): for each eta slice K the mean of
 * pi^{eta eta} / (e + P) over the cells with e + P > 0.133, summed over i, then j */
void orc_pizz_o_ep(const orc_params *p, const real *ev, const real *pi, real *out) {
    const int64_t NY = p->ny, NZ = p->nz;
    for (int64_t K = 0; K < NZ; K++) {
        real ratio = RC(0.0f);
        int num_of_cells = 0;
        for (int64_t i = 0; i < p->nx; i++)
            for (int64_t j = 0; j < NY; j++) {
                const int64_t I = i * NY * NZ + j * NZ + K;
                const real ep = ev[4 * I] + eos_P(p, ev[4 * I]);
                if (ep > RC(0.133f)) {
                    ratio += DIV(pi[10 * I + 9], ep);
                    num_of_cells++;
                }
            }
        out[K] = DIV(ratio, (real)num_of_cells);
    }
}

/* ------------------------------------------------------------------------- */
/* Pencil iteration                                                           */
/* ------------------------------------------------------------------------- */
/* Cells of the pencil along `axis` through (a, b) of the two other axes:
 * returns the base index and stride so that cell t of the pencil is base + t*stride. */
static void pencil(const orc_params *p, int axis, int64_t q, int64_t *base, int64_t *stride, int *len) {
    int64_t NY = p->ny, NZ = p->nz;
    if (axis == 0) {        /* q runs over (j,k) */
        *base = q; *stride = NY * NZ; *len = p->nx;
    } else if (axis == 1) { /* q runs over (i,k) */
        int64_t i = q / NZ, k = q % NZ;
        *base = i * NY * NZ + k; *stride = NZ; *len = p->ny;
    } else {                /* q runs over (i,j) */
        *base = q * NZ; *stride = 1; *len = p->nz;
    }
}
static int64_t npencils(const orc_params *p, int axis) {
    if (axis == 0) return (int64_t)p->ny * p->nz;
    if (axis == 1) return (int64_t)p->nx * p->nz;
    return (int64_t)p->nx * p->ny;
}
/* 2-ghost constant extrapolation at both ends of a pencil (kernel_ideal.cl:57-62) */
static int64_t ghost(int64_t base, int64_t stride, int len, int t) {
    if (t < 0) t = 0;
    if (t > len - 1) t = len - 1;
    return base + (int64_t)t * stride;
}

static void set_threads(const orc_params *p) {
#ifdef _OPENMP
    if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#else
    (void)p;
#endif
}

/* ------------------------------------------------------------------------- */
/* Ideal hydro kernels (kernel_ideal.cl)                                      */
/* ------------------------------------------------------------------------- */

/* kt_src_christoffel, kernel_ideal.cl:3-36 */
void orc_ideal_christoffel(const orc_params *p, real *Src, const real *ev, real tau, int step) {
    int64_t n = ncells(p);
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < n; I++) {
        real *s = Src + 4 * I;
        if (step == 1) { s[0] = s[1] = s[2] = s[3] = RC(0.0f); }
        const real *e = ev + 4 * I;
        real ed = e[0], vx = e[1], vy = e[2], vz = e[3];
        real pressure = eos_P(p, ed);
        real u0 = gamma3(vx, vy, vz);
        real Tzz = (ed + pressure) * u0 * u0 * vz * vz + pressure;
        real Ttz = (ed + pressure) * u0 * u0 * vz;
        if (!p->riemann_test) {
            s[0] = s[0] - Tzz;
            s[1] = s[1] - RC(0.0f);
            s[2] = s[2] - RC(0.0f);
            s[3] = s[3] - Ttz;
        }
    }
}

/* kt1d, helper.h:179-242: H(i+1/2) - H(i-1/2) for 5 stencil cells c[0..4] */
static void kt1d(const orc_params *p, const real *c[5], int along, real out[4]) {
    real pr_im1 = eos_P(p, c[1][0]), pr_i = eos_P(p, c[2][0]);
    real pr_ip1 = eos_P(p, c[3][0]), pr_ip2 = eos_P(p, c[4][0]);
    real Tm1[4], Ti[4], Tp1[4], Tp2[4], Tm2[4];
    t0m(c[1], pr_im1, Tm1); t0m(c[2], pr_i, Ti); t0m(c[3], pr_ip1, Tp1); t0m(c[4], pr_ip2, Tp2);
    real DA0[4], DA1[4], AL[4], AR[4], src[4];
    for (int q = 0; q < 4; q++) {
        DA0[q] = slope(Tm1[q], Ti[q], Tp1[q]);
        DA1[q] = slope(Ti[q], Tp1[q], Tp2[q]);
        AL[q] = Ti[q] + RC(0.5f) * DA0[q];
        AR[q] = Tp1[q] - RC(0.5f) * DA1[q];
    }
    real pr_half = RC(0.5f) * (pr_ip1 + pr_i);
    real vi_half = RC(0.5f) * (c[2][1 + along] + c[3][1 + along]);
    real evh[4];
    for (int q = 0; q < 4; q++) evh[q] = RC(0.5f) * (c[2][q] + c[3][q]);
    real lam = max_prop_speed(p, evh, vi_half);
    for (int q = 0; q < 4; q++) {
        real g0 = q == 0 ? RC(1.0f) : RC(0.0f);
        real ga = q == along + 1 ? RC(-1.0f) : RC(0.0f);
        real Jp = (AR[q] + pr_half * g0) * vi_half - pr_half * ga;
        real Jm = (AL[q] + pr_half * g0) * vi_half - pr_half * ga;
        src[q] = RC(0.5f) * (Jp + Jm) - RC(0.5f) * lam * (AR[q] - AL[q]);
    }
    real pr_im2 = eos_P(p, c[0][0]);
    t0m(c[0], pr_im2, Tm2);
    for (int q = 0; q < 4; q++) {
        DA1[q] = DA0[q];
        DA0[q] = slope(Tm2[q], Tm1[q], Ti[q]);
        AL[q] = Tm1[q] + RC(0.5f) * DA0[q];
        AR[q] = Ti[q] - RC(0.5f) * DA1[q];
    }
    pr_half = RC(0.5f) * (pr_im1 + pr_i);
    vi_half = RC(0.5f) * (c[1][1 + along] + c[2][1 + along]);
    for (int q = 0; q < 4; q++) evh[q] = RC(0.5f) * (c[2][q] + c[1][q]);
    lam = max_prop_speed(p, evh, vi_half);
    for (int q = 0; q < 4; q++) {
        real g0 = q == 0 ? RC(1.0f) : RC(0.0f);
        real ga = q == along + 1 ? RC(-1.0f) : RC(0.0f);
        real Jp = (AR[q] + pr_half * g0) * vi_half - pr_half * ga;
        real Jm = (AL[q] + pr_half * g0) * vi_half - pr_half * ga;
        src[q] -= RC(0.5f) * (Jp + Jm) - RC(0.5f) * lam * (AR[q] - AL[q]);
        out[q] = src[q];
    }
}

/* kt_src_alongx/y/z, kernel_ideal.cl:39-160 */
void orc_ideal_along(const orc_params *p, int axis, real *Src, const real *ev, real tau) {
    int64_t np_ = npencils(p, axis);
    const real DX = RC(axis == 0 ? p->dx : (axis == 1 ? p->dy : p->dz));
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < np_; q++) {
        int64_t base, stride; int len;
        pencil(p, axis, q, &base, &stride, &len);
        for (int t = 0; t < len; t++) {
            const real *c[5];
            for (int s = 0; s < 5; s++) c[s] = ev + 4 * ghost(base, stride, len, t + s - 2);
            real k[4];
            kt1d(p, c, axis, k);
            real *S = Src + 4 * (base + (int64_t)t * stride);
            for (int m = 0; m < 4; m++) {
                if (p->riemann_test) S[m] = S[m] - DIV(k[m], DX);
                else if (axis < 2) S[m] = S[m] - DIV(tau * k[m], DX);
                else S[m] = S[m] - DIV(tau * k[m], (p->plant & ORC_PLANT_ETA_DZ) ? DX : tau * DX);
            }
        }
    }
}

/* update_ev, kernel_ideal.cl:165-231 */
void orc_ideal_update_ev(const orc_params *p, real *evnew, const real *ev1, const real *Src,
                         real tau, int step, int32_t *iters) {
    int64_t n = ncells(p);
    const real DT = RC(p->dt);
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < n; I++) {
        const real *e = ev1 + 4 * I;
        real ed = e[0], vx = e[1], vy = e[2], vz = e[3];
        real pressure = eos_P(p, ed);
        real u0 = gamma3(vx, vy, vz);
        real umu[4] = {u0 * RC(1.0f), u0 * vx, u0 * vy, u0 * vz};
        real old_time = tau - (step - 1) * DT;
        real T0m[4];
        for (int q = 0; q < 4; q++) {
            real g0 = q == 0 ? RC(1.0f) : RC(0.0f);
            real t = (ed + pressure) * u0 * umu[q] - pressure * g0;
            T0m[q] = p->riemann_test ? t : t * old_time;
        }
        for (int q = 0; q < 4; q++) T0m[q] = T0m[q] + DIV(Src[4 * I + q] * DT, (real)step);
        real T00, T01, T02, T03;
        if (p->riemann_test) {
            T00 = r_max(ACU, T0m[0]);
            T01 = ((real)fabs((double)T0m[1]) < ACU) ? RC(0.0f) : T0m[1];
            T02 = ((real)fabs((double)T0m[2]) < ACU) ? RC(0.0f) : T0m[2];
            T03 = ((real)fabs((double)T0m[3]) < ACU) ? RC(0.0f) : T0m[3];
        } else {
            T00 = DIV(r_max(ACU, T0m[0]), tau);
            T01 = ((real)fabs((double)T0m[1]) < ACU) ? RC(0.0f) : DIV(T0m[1], tau);
            T02 = ((real)fabs((double)T0m[2]) < ACU) ? RC(0.0f) : DIV(T0m[2], tau);
            T03 = ((real)fabs((double)T0m[3]) < ACU) ? RC(0.0f) : DIV(T0m[3], tau);
        }
        real M = (real)sqrt((double)(T01 * T01 + T02 * T02 + T03 * T03));
        real SC = RC(0.999f);
        if (M > T00) {
            T01 *= DIV(SC * T00, M);
            T02 *= DIV(SC * T00, M);
            T03 *= DIV(SC * T00, M);
            M = SC * T00;
        }
        int it = 0;
        real ed_find = solve_ed(p, T00, M, RC(0.0f), 0, &it);
        if (iters) iters[I] = it;
        ed_find = r_max(RC(0.0f), ed_find);
        real pr = eos_P(p, ed_find);
        real epv = r_max(ACU, T00 + pr);
        real *o = evnew + 4 * I;
        o[0] = ed_find; o[1] = DIV(T01, epv); o[2] = DIV(T02, epv); o[3] = DIV(T03, epv);
    }
}

/* ------------------------------------------------------------------------- */
/* Viscous KT kernels (kernel_visc.cl), with the optional bulk Pi extension   */
/* ------------------------------------------------------------------------- */
static real bulk_at(const real *bulk, int64_t I) { return bulk ? bulk[I] : RC(0.0f); }

/* kt_src_christoffel (visc), kernel_visc.cl:11-53 */
void orc_visc_christoffel(const orc_params *p, real *Src, const real *ev, const real *pi,
                          const real *bulk, real tau, int step) {
    int64_t n = ncells(p);
    set_threads(p);
    (void)tau;
    #pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < n; I++) {
        real *s = Src + 4 * I;
        if (step == 1) { s[0] = s[1] = s[2] = s[3] = RC(0.0f); }
        const real *e = ev + 4 * I;
        real ed = e[0], vx = e[1], vy = e[2], vz = e[3];
        real u0 = gamma3(vx, vy, vz);
        real bulkpi = bulk_at(bulk, I);
        real pressure = eos_P(p, ed) + bulkpi;
        real Tzz = (ed + pressure) * u0 * u0 * vz * vz + pressure + pi[10 * I + 9];
        real Ttz = (ed + pressure) * u0 * u0 * vz + pi[10 * I + 3];
        s[0] = s[0] - Tzz;
        s[1] = s[1] - RC(0.0f);
        s[2] = s[2] - RC(0.0f);
        s[3] = s[3] - Ttz;
        if (p->riemann_test) {
            real Ttt = (ed + pressure) * u0 * u0 - pressure;
            real Ttx = (ed + pressure) * u0 * u0 * vx;
            real Tty = (ed + pressure) * u0 * u0 * vy;
            s[0] = s[0] + (Tzz + Ttt);
            s[1] = s[1] + Ttx;
            s[2] = s[2] + Tty;
            s[3] = s[3] + RC(3.0f) * Ttz;
        }
    }
}

/* row mu of pi for the flux along `along`: pi^{a+1, mu} (kernel_visc.cl:151-154) */
static const int PI_ROW[4][4] = {
    {0, 1, 2, 3},   /* pi^{0 mu} */
    {1, 4, 5, 6},   /* pi^{1 mu} */
    {2, 5, 7, 8},   /* pi^{2 mu} */
    {3, 6, 8, 9},   /* pi^{3 mu} */
};

/* kt1d_visc, kernel_visc.cl:60-127 */
static void kt1d_visc(const orc_params *p, const real *ev[5], const real *pi[5], const real *bk[5],
                      real vip_half, real vim_half, real tau, int along, real out[4]) {
    real pr[5], Q[5][4], pim0[5][4], pimi[5][4];
    for (int i = 0; i < 5; i++) {
        pr[i] = eos_P(p, ev[i][0]) + (bk[i] ? *bk[i] : RC(0.0f));
        real T[4];
        t0m(ev[i], pr[i], T);
        for (int q = 0; q < 4; q++) {
            pim0[i][q] = pi[i][PI_ROW[0][q]];
            pimi[i][q] = pi[i][PI_ROW[along + 1][q]];
            Q[i][q] = tau * (T[q] + pim0[i][q]);
        }
    }
    real DA0[4], DA1[4], AL[4], AR[4], src[4];
    for (int q = 0; q < 4; q++) {
        DA0[q] = slope(Q[1][q], Q[2][q], Q[3][q]);
        DA1[q] = slope(Q[2][q], Q[3][q], Q[4][q]);
        AL[q] = Q[2][q] + RC(0.5f) * DA0[q];
        AR[q] = Q[3][q] - RC(0.5f) * DA1[q];
    }
    real pr_half = RC(0.5f) * (pr[2] + pr[3]);
    real evh[4];
    for (int q = 0; q < 4; q++) evh[q] = RC(0.5f) * (ev[2][q] + ev[3][q]);
    real lam = max_prop_speed(p, evh, vip_half);
    for (int q = 0; q < 4; q++) {
        real g0 = q == 0 ? RC(1.0f) : RC(0.0f);
        real ga = q == along + 1 ? RC(-1.0f) : RC(0.0f);
        real pim0_half = RC(0.5f) * (pim0[2][q] + pim0[3][q]);
        real pimi_half = RC(0.5f) * (pimi[2][q] + pimi[3][q]);
        real Jp = (AR[q] + pr_half * tau * g0 - tau * pim0_half) * vip_half - pr_half * tau * ga + tau * pimi_half;
        real Jm = (AL[q] + pr_half * tau * g0 - tau * pim0_half) * vip_half - pr_half * tau * ga + tau * pimi_half;
        src[q] = RC(0.5f) * (Jp + Jm) - RC(0.5f) * lam * (AR[q] - AL[q]);
    }
    for (int q = 0; q < 4; q++) {
        DA1[q] = DA0[q];
        DA0[q] = slope(Q[0][q], Q[1][q], Q[2][q]);
        AL[q] = Q[1][q] + RC(0.5f) * DA0[q];
        AR[q] = Q[2][q] - RC(0.5f) * DA1[q];
    }
    pr_half = RC(0.5f) * (pr[1] + pr[2]);
    for (int q = 0; q < 4; q++) evh[q] = RC(0.5f) * (ev[2][q] + ev[1][q]);
    lam = max_prop_speed(p, evh, vim_half);
    for (int q = 0; q < 4; q++) {
        real g0 = q == 0 ? RC(1.0f) : RC(0.0f);
        real ga = q == along + 1 ? RC(-1.0f) : RC(0.0f);
        real pim0_half = RC(0.5f) * (pim0[1][q] + pim0[2][q]);
        real pimi_half = RC(0.5f) * (pimi[1][q] + pimi[2][q]);
        real Jp = (AR[q] + pr_half * tau * g0 - tau * pim0_half) * vim_half - pr_half * tau * ga + tau * pimi_half;
        real Jm = (AL[q] + pr_half * tau * g0 - tau * pim0_half) * vim_half - pr_half * tau * ga + tau * pimi_half;
        src[q] -= RC(0.5f) * (Jp + Jm) - RC(0.5f) * lam * (AR[q] - AL[q]);
        out[q] = src[q];
    }
}

/* kt_src_alongx/y/z (visc), kernel_visc.cl:131-318 */
void orc_visc_along(const orc_params *p, int axis, real *Src, const real *ev, const real *pi,
                    const real *bulk, real tau) {
    int64_t np_ = npencils(p, axis);
    const real DX = RC(axis == 0 ? p->dx : (axis == 1 ? p->dy : p->dz));
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < np_; q++) {
        int64_t base, stride; int len;
        pencil(p, axis, q, &base, &stride, &len);
        for (int t = 0; t < len; t++) {
            const real *c[5], *pc[5], *bc[5];
            for (int s = 0; s < 5; s++) {
                int64_t J = ghost(base, stride, len, t + s - 2);
                c[s] = ev + 4 * J; pc[s] = pi + 10 * J; bc[s] = bulk ? bulk + J : NULL;
            }
            real vip_half = RC(0.5f) * (c[2][1 + axis] + c[3][1 + axis]);
            real vim_half = RC(0.5f) * (c[1][1 + axis] + c[2][1 + axis]);
            real k[4];
            kt1d_visc(p, c, pc, bc, vip_half, vim_half, tau, axis, k);
            real *S = Src + 4 * (base + (int64_t)t * stride);
            for (int m = 0; m < 4; m++) {
                if (axis < 2) S[m] = S[m] - DIV(k[m], DX);
                else S[m] = S[m] - DIV(k[m], (p->plant & ORC_PLANT_ETA_DZ) ? DX : tau * DX);
            }
        }
    }
}

/* update_ev (visc) from the clamp |T^ti| <= 0.999 T^tt on, kernel_visc.cl:365-392: the
 * root solve (from *v0 when given) and the velocities, written to o[0..3] */
static void visc_finish(const orc_params *p, real T00, real T01, real T02, real T03, real bnew, const real *v0,
                        int *it, real *o) {
    real M = (real)sqrt((double)(T01 * T01 + T02 * T02 + T03 * T03));
    real SC = RC(0.999f);
    if (M > T00) {
        T01 *= DIV(SC * T00, M);
        T02 *= DIV(SC * T00, M);
        T03 *= DIV(SC * T00, M);
        M = SC * T00;
    }
    /* the FP64 root of fast cells (fp64_root) comes before a seed */
    const int fast = p->fp64_root && M > RC(p->fp64_root_vmin) * T00;
    real ed_find = (v0 && !fast) ? root_newton_v0(p, T00, M, bnew, p->bulk_on, *v0, it)
                                 : solve_ed(p, T00, M, bnew, p->bulk_on, it);
    if (ed_find < ACU) { T01 = RC(0.0f); T02 = RC(0.0f); T03 = RC(0.0f); }
    ed_find = r_max(RC(0.0f), ed_find);
    real pr = eos_P(p, ed_find) + bnew;
    if (p->bulk_on) pr = r_max(pr, RC(0.0f));
    real epv = r_max(ACU, T00 + pr);
    o[0] = ed_find; o[1] = DIV(T01, epv); o[2] = DIV(T02, epv); o[3] = DIV(T03, epv);
}

/* update_ev (visc), kernel_visc.cl:322-393; bulk1/bulk2 are Pi at the start point
 * and the freshly updated Pi (NULL without the bulk extension). */
void orc_visc_update_ev(const orc_params *p, real *evnew, const real *ev1, const real *pi1,
                        const real *pi2, const real *bulk1, const real *bulk2, const real *Src,
                        real tau, int step, int32_t *iters) {
    int64_t n = ncells(p);
    const real DT = RC(p->dt);
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < n; I++) {
        const real *e = ev1 + 4 * I;
        real ed = e[0], vx = e[1], vy = e[2], vz = e[3];
        real pressure = eos_P(p, ed) + bulk_at(bulk1, I);
        real u0 = gamma3(vx, vy, vz);
        real umu[4] = {u0 * RC(1.0f), u0 * vx, u0 * vy, u0 * vz};
        real old_time = tau - (step - 1) * DT;
        real T0m[4];
        for (int q = 0; q < 4; q++) {
            real g0 = q == 0 ? RC(1.0f) : RC(0.0f);
            T0m[q] = ((ed + pressure) * u0 * umu[q] - pressure * g0 + pi1[10 * I + q]) * old_time;
        }
        for (int q = 0; q < 4; q++)
            T0m[q] = T0m[q] + DIV(Src[4 * I + q] * DT, (real)step) - pi2[10 * I + q] * tau;
        real T00 = DIV(r_max(ACU, T0m[0]), tau);
        real T01 = ((real)fabs((double)T0m[1]) < ACU) ? RC(0.0f) : DIV(T0m[1], tau);
        real T02 = ((real)fabs((double)T0m[2]) < ACU) ? RC(0.0f) : DIV(T0m[2], tau);
        real T03 = ((real)fabs((double)T0m[3]) < ACU) ? RC(0.0f) : DIV(T0m[3], tau);
        int it = 0;
        visc_finish(p, T00, T01, T02, T03, bulk_at(bulk2, I), NULL, &it, evnew + 4 * I);
        if (iters) iters[I] = it;
    }
}

/* The end of update_ev (visc) for given conserved inputs (T^tt, T^ti as update_ev has them
 * before the clamp, 4 per cell) and Pi_new (NULL: no bulk): the split form of the stage.
 * v0 (NULL: the reference's start) seeds the root solve, clamped to [0, 0.9999]; cells that
 * take the FP64 root (fp64_root, M > vmin T00) take it with or without a seed. */
void orc_visc_root_from(const orc_params *p, const real *cons, const real *bulk, const real *v0, int64_t n,
                        real *ev_out, int32_t *iters) {
    for (int64_t I = 0; I < n; I++) {
        const real *c = cons + 4 * I;
        int it = 0;
        real s, *sp = NULL;
        if (v0) {
            s = v0[I] > RC(0.0f) ? v0[I] : RC(0.0f);
            s = s < RC(0.9999f) ? s : RC(0.9999f);
            sp = &s;
        }
        visc_finish(p, c[0], c[1], c[2], c[3], bulk ? bulk[I] : RC(0.0f), sp, &it, ev_out + 4 * I);
        if (iters) iters[I] = it;
    }
}

/* ------------------------------------------------------------------------- */
/* Israel-Stewart kernels (kernel_IS.cl)                                      */
/* ------------------------------------------------------------------------- */
static const real GMN_DIAG[4] = {RC(1.0f), RC(-1.0f), RC(-1.0f), RC(-1.0f)};
static int IDX(int i, int j) {
    return (i < j) ? ((7 * i + 2 * j - i * i) / 2) : ((7 * j + 2 * i - j * j) / 2);
}

/* visc_initialize, kernel_IS.cl:10-36 */
void orc_visc_initialize(const orc_params *p, real *pi1, real *goodcell, real *udiff,
                         const real *ev, real tau) {
    int64_t n = ncells(p);
    for (int64_t I = 0; I < n; I++) {
        real cpTs[4];
        eos4(p, ev[4 * I], cpTs);
        real etav = etaos(p, cpTs[2]) * cpTs[3] * HBARC;
        real tmp = DIV(RC(2.0f) / RC(3.0f) * etav, tau);
        pi1[10 * I + 4] = tmp;
        pi1[10 * I + 7] = tmp;
        pi1[10 * I + 9] = RC(-2.0f) * tmp;
        udiff[4 * I + 0] = udiff[4 * I + 1] = udiff[4 * I + 2] = udiff[4 * I + 3] = RC(0.0f);
        goodcell[I] = RC(1.0f);
    }
}

/* visc_src_christoffel, kernel_IS.cl:41-67 */
void orc_is_christoffel(const orc_params *p, real *Src, const real *pi1, const real *ev,
                        real tau, int step) {
    int64_t n = ncells(p);
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < n; I++) {
        real *s = Src + 10 * I;
        const real *q = pi1 + 10 * I;
        if (step == 1) for (int m = 0; m < 10; m++) s[m] = RC(0.0f);
        const real *e = ev + 4 * I;
        real uz = e[3] * gamma3(e[1], e[2], e[3]);
        s[0] -= DIV(RC(2.0f) * uz * q[3], tau);
        s[1] -= DIV(uz * q[6], tau);
        s[2] -= DIV(uz * q[8], tau);
        s[3] -= DIV(uz * (q[0] + q[9]), tau);
        s[6] -= DIV(uz * q[1], tau);
        s[8] -= DIV(uz * q[2], tau);
        s[9] -= DIV(RC(2.0f) * uz * q[3], tau);
    }
}

/* kt1d_real, helper.h:249-285 */
static real kt1d_real(real Q_im2, real Q_im1, real Q_i, real Q_ip1, real Q_ip2,
                      real v_mh, real v_ph, real lam_mh, real lam_ph) {
    real DA0 = slope(Q_im1, Q_i, Q_ip1);
    real DA1 = slope(Q_i, Q_ip1, Q_ip2);
    real AL = Q_i + RC(0.5f) * DA0;
    real AR = Q_ip1 - RC(0.5f) * DA1;
    real Jp = AR * v_ph, Jm = AL * v_ph;
    real src = RC(0.5f) * (Jp + Jm) - RC(0.5f) * lam_ph * (AR - AL);
    DA1 = DA0;
    DA0 = slope(Q_im2, Q_im1, Q_i);
    AL = Q_im1 + RC(0.5f) * DA0;
    AR = Q_i - RC(0.5f) * DA1;
    Jp = AR * v_mh; Jm = AL * v_mh;
    src -= RC(0.5f) * (Jp + Jm) - RC(0.5f) * lam_mh * (AR - AL);
    return src;
}

/* visc_src_alongx/y/z, kernel_IS.cl:74-318.  ud receives the limited gradient of u^mu
 * along the axis (plus the Christoffel term for eta); bulk (optional) is evolved as an
 * 11th component with the same flux. */
void orc_is_along(const orc_params *p, int axis, real *Src, real *ud, const real *pi1,
                  const real *ev, real *bulk_src, const real *bulk, real tau) {
    int64_t np_ = npencils(p, axis);
    const real DX = RC(axis == 0 ? p->dx : (axis == 1 ? p->dy : p->dz));
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < np_; q++) {
        int64_t base, stride; int len;
        pencil(p, axis, q, &base, &stride, &len);
        for (int t = 0; t < len; t++) {
            int64_t J[5];
            for (int s = 0; s < 5; s++) J[s] = ghost(base, stride, len, t + s - 2);
            const real *e[5];
            real u0[5];
            for (int s = 0; s < 5; s++) { e[s] = ev + 4 * J[s]; u0[s] = gamma_ev(e[s]); }
            real v_mh = RC(0.5f) * (e[1][1 + axis] + e[2][1 + axis]);
            real v_ph = RC(0.5f) * (e[3][1 + axis] + e[2][1 + axis]);
            real evm[4], evp[4];
            for (int c = 0; c < 4; c++) {
                evm[c] = RC(0.5f) * (e[1][c] + e[2][c]);
                evp[c] = RC(0.5f) * (e[3][c] + e[2][c]);
            }
            real lam_mh = max_prop_speed(p, evm, v_mh);
            real lam_ph = max_prop_speed(p, evp, v_ph);
            int64_t IND = base + (int64_t)t * stride;
            real um1[4], ui[4], up1[4];
            umu4(e[1], um1); umu4(e[2], ui); umu4(e[3], up1);
            for (int c = 0; c < 4; c++) {
                real d = slope(um1[c], ui[c], up1[c]);
                if (axis < 2) ud[4 * IND + c] = DIV(d, DX);
                else {
                    real chr = DIV((c == 0 ? e[2][3] : (c == 3 ? RC(1.0f) : RC(0.0f))) * u0[2], tau);
                    if (p->plant & ORC_PLANT_UDZ_NO_CHRISTOFFEL) chr = RC(0.0f);
                    ud[4 * IND + c] = DIV(d, (p->plant & ORC_PLANT_ETA_DZ) ? DX : DX * tau) + chr;
                }
            }
            for (int mn = 0; mn < 10; mn++) {
                real k = kt1d_real(u0[0] * pi1[10 * J[0] + mn], u0[1] * pi1[10 * J[1] + mn],
                                   u0[2] * pi1[10 * J[2] + mn], u0[3] * pi1[10 * J[3] + mn],
                                   u0[4] * pi1[10 * J[4] + mn], v_mh, v_ph, lam_mh, lam_ph);
                if (axis < 2) Src[10 * IND + mn] = Src[10 * IND + mn] - DIV(k, DX);
                else Src[10 * IND + mn] = Src[10 * IND + mn] - DIV(k, (p->plant & ORC_PLANT_ETA_DZ) ? DX : tau * DX);
            }
            if (bulk && bulk_src) {
                real k = kt1d_real(u0[0] * bulk[J[0]], u0[1] * bulk[J[1]], u0[2] * bulk[J[2]],
                                   u0[3] * bulk[J[3]], u0[4] * bulk[J[4]], v_mh, v_ph, lam_mh, lam_ph);
                if (axis < 2) bulk_src[IND] = bulk_src[IND] - DIV(k, DX);
                else bulk_src[IND] = bulk_src[IND] - DIV(k, (p->plant & ORC_PLANT_ETA_DZ) ? DX : tau * DX);
            }
        }
    }
}

/* PiPi self coupling used by the GUBSER_VISC_TEST branch, kernel_IS.cl:322-347 */
static real dot4(const real *a, const real *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}
static real pipi(int lam, int mu, int nu, const real *pim, const real *u) {
    real A[4] = {pim[IDX(0, lam)], -pim[IDX(1, lam)], -pim[IDX(2, lam)], -pim[IDX(3, lam)]};
    real D[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) D[a][b] = (a == b ? GMN_DIAG[a] : RC(0.0f)) - u[a] * u[b];
    real first = dot4(D[mu], A) * dot4(D[nu], A);
    real temp[4] = {dot4(D[0], A), dot4(D[1], A), dot4(D[2], A), dot4(D[3], A)};
    real second = RC(-1.0f) / RC(3.0f) * ((mu == nu ? GMN_DIAG[mu] : RC(0.0f)) - u[mu] * u[nu]) * dot4(temp, A);
    return first + second;
}

/* Vorticity couplings of update_pimn (kernel_IS.cl:349-470): Omega^{mu nu} =
 * 1/2 Delta^mu_a Delta^nu_b omega^{ab} from omega = (d_t u^x - d_x u^t, d_t u^y - d_y u^t,
 * d_t u^eta - d_eta u^t, d_x u^y - d_y u^x, d_x u^eta - d_eta u^x, d_y u^eta - d_eta u^y)
 * (kernel_IS.cl:636-651), and the projected products pi^<mu_lam Omega^nu>lam and
 * Omega^<mu_lam Omega^nu>lam, each in the reference's operand order. */
static real omega_proj(int mu, int nu, const real *om, const real *u) {
    const real ul[4] = {u[0], -u[1], -u[2], -u[3]};
    real D[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) D[a][b] = (a == b ? RC(1.0f) : RC(0.0f)) - u[a] * ul[b];
    const real A[4][4] = {{RC(0.0f), om[0], om[1], om[2]}, {-om[0], RC(0.0f), om[3], om[4]},
                          {-om[1], -om[3], RC(0.0f), om[5]}, {-om[2], -om[4], -om[5], RC(0.0f)}};
    real temp[4] = {dot4(D[nu], A[0]), dot4(D[nu], A[1]), dot4(D[nu], A[2]), dot4(D[nu], A[3])};
    return RC(0.5f) * dot4(D[mu], temp);
}
static void omega_row(int lam, const real *Om, real B[4]) {
    if (lam == 0) { B[0] = RC(0.0f); B[1] = Om[0]; B[2] = Om[1]; B[3] = Om[2]; }
    else if (lam == 1) { B[0] = Om[0]; B[1] = RC(0.0f); B[2] = Om[3]; B[3] = Om[4]; }
    else if (lam == 2) { B[0] = Om[1]; B[1] = -Om[3]; B[2] = RC(0.0f); B[3] = Om[5]; }
    else { B[0] = Om[2]; B[1] = -Om[4]; B[2] = -Om[5]; B[3] = RC(0.0f); }
}
static real pi_omega(int lam, int mu, int nu, const real *pim, const real *Om, const real *u) {
    real A[4] = {pim[IDX(0, lam)], -pim[IDX(1, lam)], -pim[IDX(2, lam)], -pim[IDX(3, lam)]};
    real B[4];
    omega_row(lam, Om, B);
    real D[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) D[a][b] = (a == b ? GMN_DIAG[a] : RC(0.0f)) - u[a] * u[b];
    real first = RC(0.5f) * (dot4(D[mu], A) * dot4(D[nu], B) + dot4(D[mu], B) * dot4(D[nu], A));
    real temp[4] = {dot4(D[0], A), dot4(D[1], A), dot4(D[2], A), dot4(D[3], A)};
    real second = RC(-1.0f) / RC(3.0f) * ((mu == nu ? GMN_DIAG[mu] : RC(0.0f)) - u[mu] * u[nu]) * dot4(temp, B);
    return first + second;
}
static real omega_omega(int lam, int mu, int nu, const real *Om, const real *u) {
    real B[4];
    omega_row(lam, Om, B);
    real D[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) D[a][b] = (a == b ? GMN_DIAG[a] : RC(0.0f)) - u[a] * u[b];
    real first = dot4(D[mu], B) * dot4(D[nu], B);
    real temp[4] = {dot4(D[0], B), dot4(D[1], B), dot4(D[2], B), dot4(D[3], B)};
    real second = RC(-1.0f) / RC(3.0f) * ((mu == nu ? GMN_DIAG[mu] : RC(0.0f)) - u[mu] * u[nu]) * dot4(temp, B);
    return first + second;
}
/* the six projected vorticity components (Omega^{01}, 02, 03, 12, 13, 23) of a cell */
static void omega_cell(const real *udt, const real *ux, const real *uy, const real *uz, const real *u, real Om[6]) {
    real om[6] = {udt[1] - ux[0], udt[2] - uy[0], udt[3] - uz[0], ux[2] - uy[1], ux[3] - uz[1], uy[3] - uz[2]};
    Om[0] = omega_proj(0, 1, om, u); Om[1] = omega_proj(0, 2, om, u); Om[2] = omega_proj(0, 3, om, u);
    Om[3] = omega_proj(1, 2, om, u); Om[4] = omega_proj(1, 3, om, u); Om[5] = omega_proj(2, 3, om, u);
}
/* test entry: the coupling tensors of one state, out[0..9] = sum over lam of
 * g_lam,lam pi_omega(lam, mu, nu), out[10..19] the same of omega_omega, out[20..25] Omega */
void orc_couplings(const real *pim, const real *udt, const real *ux, const real *uy, const real *uz,
                   const real *u, real *out) {
    real Om[6];
    omega_cell(udt, ux, uy, uz, u, Om);
    for (int mu = 0; mu < 4; mu++)
        for (int nu = mu; nu < 4; nu++) {
            int mn = IDX(mu, nu);
            out[mn] = pi_omega(0, mu, nu, pim, Om, u) - pi_omega(1, mu, nu, pim, Om, u)
                      - pi_omega(2, mu, nu, pim, Om, u) - pi_omega(3, mu, nu, pim, Om, u);
            out[10 + mn] = omega_omega(0, mu, nu, Om, u) - omega_omega(1, mu, nu, Om, u)
                           - omega_omega(2, mu, nu, Om, u) - omega_omega(3, mu, nu, Om, u);
        }
    for (int q = 0; q < 6; q++) out[20 + q] = Om[q];
}

/* update_pimn, kernel_IS.cl:501-709.  With the bulk extension, Pi is relaxed to
 * -zeta*theta with tau_Pi = 3 (zeta/s) hbarc / T and source Pi*theta/3 - Pi*u^tau/tau
 * (the conservative form of tau_Pi D Pi + Pi = -zeta theta - (2/3) tau_Pi Pi theta). */
void orc_update_pimn(const orc_params *p, real *pinew, real *goodcell, const real *pi1,
                     const real *pistep, const real *ev1, const real *ev2, const real *udiff,
                     const real *udx, const real *udy, const real *udz, real *Src,
                     real *bulknew, const real *bulk1, const real *bulkstep, real *bulk_src,
                     real tau, int step) {
    int64_t n = ncells(p);
    const real DT = RC(p->dt);
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < n; I++) {
        const real *e1 = ev1 + 4 * I, *e2 = ev2 + 4 * I;
        real u_old[4], u_new[4], udt[4];
        umu4(e1, u_old); umu4(e2, u_new);
        for (int c = 0; c < 4; c++) udt[c] = DIV(u_new[c] - u_old[c], DT);
        if (step == 1) {
            for (int c = 0; c < 4; c++) {
                udt[c] = DIV(udiff[4 * I + c], DT);
                u_new[c] = u_old[c] + udiff[4 * I + c];
            }
        }
        const real *ux = udx + 4 * I, *uy = udy + 4 * I, *uz = udz + 4 * I;
        /* dalpha_u[nu] = (d_t, d_x, d_y, d_z) u^nu */
        real da[4][4];
        for (int c = 0; c < 4; c++) { da[c][0] = udt[c]; da[c][1] = ux[c]; da[c][2] = uy[c]; da[c][3] = uz[c]; }
        real DU[4], u[4], ed_step;
        const real *us = (step == 1) ? u_old : u_new;
        for (int c = 0; c < 4; c++) DU[c] = dot4(us, da[c]);
        for (int c = 0; c < 4; c++) u[c] = us[c];
        ed_step = (step == 1) ? e1[0] : e2[0];
        real theta = udt[0] + ux[1] + uy[2] + uz[3];
        real cpTs[4];
        eos4(p, ed_step, cpTs);
        real temperature = cpTs[2];
        real local_etaos = etaos(p, temperature);
        real etav = local_etaos * cpTs[3] * HBARC;
        real one_over_taupi = DIV(temperature, RC(3.0f) * r_max(ACU, local_etaos) * HBARC);
        real coef_pipi = RC(0.0f);
        if (p->gubser_visc_test) {
            real ETAOS = RC(p->etaos_ymin);
            /* pow(e, -1/4) and pow(e, 3/4) evaluated through sqrt (OpenCL's pow is a
             * 16-ulp function; the sqrt forms are reproducible on every device) */
            real q4 = (real)sqrt((double)(real)sqrt((double)ed_step));
            real etavH = DIV(ETAOS * RC(4.0f), RC(3.0f));
            real taupiH = DIV(RC(p->lam1) * RC(p->lam1) * etavH, RC(3.0f));
            one_over_taupi = rcp(taupiH * rcp(q4));
            coef_pipi = DIV(RC(p->lam1), ed_step);
            etav = DIV(ETAOS * RC(4.0f), RC(3.0f)) * ((real)sqrt((double)ed_step) * q4);
        }
        real pi2[10];
        for (int m = 0; m < 10; m++) pi2[m] = pistep[10 * I + m];
        real Om[6];
        if (p->couplings) omega_cell(udt, ux, uy, uz, u, Om);
        real max_pimn_abs = RC(0.0f);
        real out[10];
        for (int mu = 0; mu < 4; mu++) {
            for (int nu = mu; nu < 4; nu++) {
                real gmunu = (mu == nu) ? GMN_DIAG[mu] : RC(0.0f);
                real sigma = GMN_DIAG[mu] * da[nu][mu] + GMN_DIAG[nu] * da[mu][nu]
                             - (u[mu] * DU[nu] + u[nu] * DU[mu])
                             - RC(2.0f) / RC(3.0f) * (gmunu - u[mu] * u[nu]) * theta;
                int mn = IDX(mu, nu);
                if (u[0] > RC(10000.0f)) sigma = RC(0.0f);
                real piNS = etav * sigma;
                real pi_old = pi1[10 * I + mn] * u_old[0];
                real src = DIV(-pi2[mn] * theta, RC(3.0f)) - DIV(pi2[mn] * u[0], tau);
                for (int l = 0; l < 4; l++)
                    src -= (u[mu] * pi2[IDX(nu, l)] + u[nu] * pi2[IDX(mu, l)]) * DU[l] * GMN_DIAG[l];
                if (p->gubser_visc_test) {
                    src -= one_over_taupi * coef_pipi * (pipi(0, mu, nu, pi2, u) - pipi(1, mu, nu, pi2, u)
                                                         - pipi(2, mu, nu, pi2, u) - pipi(3, mu, nu, pi2, u));
                }
                if (p->couplings & 1)
                    src += RC(2.0f) * (pi_omega(0, mu, nu, pi2, Om, u) - pi_omega(1, mu, nu, pi2, Om, u)
                                       - pi_omega(2, mu, nu, pi2, Om, u) - pi_omega(3, mu, nu, pi2, Om, u));
                if (p->couplings & 2)
                    src -= RC(1.0f) * one_over_taupi * (omega_omega(0, mu, nu, Om, u) - omega_omega(1, mu, nu, Om, u)
                                                        - omega_omega(2, mu, nu, Om, u) - omega_omega(3, mu, nu, Om, u));
                Src[10 * I + mn] += src;
                pi_old = DIV(pi_old + DIV(Src[10 * I + mn] * DT, (real)step) + DT * one_over_taupi * piNS,
                             u_new[0] + DT * one_over_taupi);
                if ((real)fabs((double)pi_old) > max_pimn_abs) max_pimn_abs = (real)fabs((double)pi_old);
                out[mn] = pi_old;
            }
        }
        real bout = RC(0.0f);
        if (p->bulk_on && bulknew) {
            real zs = zetaos(p, temperature);
            real zetav = zs * cpTs[3] * HBARC;
            real one_over_taubulk = DIV(temperature, RC(3.0f) * r_max(ACU, zs) * HBARC);
            real b2 = bulkstep[I];
            real bsrc = DIV(b2 * theta, RC(3.0f)) - DIV(b2 * u[0], tau);
            bulk_src[I] += bsrc;
            real b_old = bulk1[I] * u_old[0];
            bout = DIV(b_old + DIV(bulk_src[I] * DT, (real)step) - DT * one_over_taubulk * zetav * theta,
                       u_new[0] + DT * one_over_taubulk);
            /* regulation (extension): |Pi| <= P/2, Pi = 0 where e < 1e-3 GeV/fm^3 */
            real hp = RC(0.5f) * eos_P(p, e2[0]);
            bout = r_min(bout, hp);
            bout = r_max(bout, -hp);
            if (e2[0] < RC(1e-3f)) bout = RC(0.0f);
            if ((real)fabs((double)bout) > max_pimn_abs) max_pimn_abs = (real)fabs((double)bout);
        }
        real pre = eos_P(p, e2[0]);
        /* bulk extension with zeta/s > 0: pi and Pi scaled together so that the largest
         * |component| is at most e + P (no reference counterpart) */
        if (p->bulk_on && bulknew && p->zetaos_max > 0.0) {
            real lim = e2[0] + pre;
            if (max_pimn_abs > lim) {
                real sc = DIV(lim, max_pimn_abs);
                for (int m = 0; m < 10; m++) out[m] = out[m] * sc;
                bout = bout * sc;
                max_pimn_abs = lim;
            }
        }
        real T00 = (e2[0] + pre) * u_new[0] * u_new[0] - pre;
        goodcell[I] = RC(1.0f);
        if (max_pimn_abs > T00) {
            for (int m = 0; m < 10; m++) out[m] = RC(0.0f);
            bout = RC(0.0f);
            goodcell[I] = RC(0.0f);
        }
        for (int m = 0; m < 10; m++) pinew[10 * I + m] = out[m];
        if (p->bulk_on && bulknew) bulknew[I] = bout;
    }
}

/* get_udiff, kernel_IS.cl:714-724 */
void orc_get_udiff(const orc_params *p, real *udiff, const real *ev0, const real *ev1) {
    int64_t n = ncells(p);
    for (int64_t I = 0; I < n; I++) {
        real a[4], b[4];
        umu4(ev0 + 4 * I, a); umu4(ev1 + 4 * I, b);
        for (int c = 0; c < 4; c++) udiff[4 * I + c] = b[c] - a[c];
    }
}

/* ------------------------------------------------------------------------- */
/* Driver-level sequencing (ideal.py / visc.py)                               */
/* ------------------------------------------------------------------------- */
typedef struct orc_state {
    real *ev[3];      /* d_ev[0..2] */
    real *pi[3];      /* d_pi[0..2] */
    real *bulk[3];    /* bulk Pi [0..2] (extension; NULL when off) */
    real *Src;        /* d_Src (ideal.d_Src) */
    real *IS_src;     /* d_IS_src */
    real *bulk_src;   /* IS source of Pi (extension) */
    real *udx, *udy, *udz, *udiff, *goodcell;
    int32_t *iters;   /* optional: root-solve iteration counts of the last update_ev */
} orc_state;

/* jet_eloss_src, kernel_jet_eloss.cl:8-69 (called at ideal.py:240-246, visc.py:265-271 after
 * the KT sources): Src += (tau, tau, tau, tau^2) N exp(-|x - X(tau)|^2 / (2 sigma^2))
 * (dE/dx)(1, cos phi, sin phi, 0), sigma = 0.3 fm, N = 1 / ((2 pi)^(3/2) sigma^3 tau),
 * X(tau) = X(start cell) + (tau - TAU0)(cos phi, sin phi, 0), dE/dx = 14 (T/0.3)^3 with T at the
 * cell of X(tau) in ev; nothing when X(tau) is off the grid.  Divisions by constants through the
 * reciprocal; exp and the cube in double, rounded once. */
void orc_jet_src(const orc_params *p, real *Src, const real *ev, real tau) {
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const int64_t nyz = (int64_t)ny * nz, st = p->jet_start;
    const int i0 = (int)(st / nyz), j0 = (int)((st - i0 * nyz) / nz), k0 = (int)(st - i0 * nyz - (int64_t)j0 * nz);
    const real c = (real)cos((double)p->jet_direction), sn = (real)sin((double)p->jet_direction);
    const real c3 = (real)(float)pow(sqrt(2.0 * M_PI), 3.0);
    const real sig = RC(0.3f), dx = RC(p->dx), dy = RC(p->dy), dz = RC(p->dz);
    const real dtau = tau - RC(p->tau0);
    const real px = (real)(float)((float)(i0 - nx / 2) * p->dx) + c * dtau;
    const real py = (real)(float)((float)(j0 - ny / 2) * p->dy) + sn * dtau;
    const real pz = (real)(float)((float)(k0 - nz / 2) * p->dz) + RC(0.0f) * dtau;
    const int li = (int)(px * (RC(1.0f) / dx) + (real)(nx / 2));
    const int lj = (int)(py * (RC(1.0f) / dy) + (real)(ny / 2));
    const int lk = (int)(pz * (RC(1.0f) / dz) + (real)(nz / 2));
    if (!(li >= 0 && li < nx && lj >= 0 && lj < ny && lk >= 0 && lk < nz)) return;
    const real e = ev[4 * ((int64_t)li * nyz + (int64_t)lj * nz + lk)];
    const double tr = (double)(eos_T(p, e) * (RC(1.0f) / RC(0.3f)));
    const real dEdx = RC(14.0f) * (real)(tr * tr * tr);
    const real dE[4] = {dEdx, dEdx * c, dEdx * sn, RC(0.0f)};
    const real norm = (real)(1.0 / (double)(c3 * sig * sig * sig * tau));
    const real r2 = RC(1.0f) / (RC(2.0f) * sig * sig);
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < ncells(p); I++) {
        const int i = (int)(I / nyz), j = (int)((I - i * nyz) / nz), k = (int)(I - i * nyz - (int64_t)j * nz);
        const real x = (real)(i - nx / 2) * dx, y = (real)(j - ny / 2) * dy, z = (real)(k - nz / 2) * dz;
        const real g = (-(x - px)) * (x - px) * r2 - (y - py) * (y - py) * r2 - (z - pz) * (z - pz) * r2;
        const real f = norm * (real)exp((double)g);
        const real coef[4] = {tau * f, tau * f, tau * f, (tau * tau) * f};
        for (int q = 0; q < 4; q++) Src[4 * I + q] = Src[4 * I + q] + coef[q] * dE[q];
    }
}

/* CLIdeal.stepUpdate(step), ideal.py:211-255 */
void orc_ideal_stage(const orc_params *p, orc_state *s, int step, real tau) {
    orc_ideal_christoffel(p, s->Src, s->ev[step], tau, step);
    for (int a = 0; a < 3; a++) orc_ideal_along(p, a, s->Src, s->ev[step], tau);
    if (p->jet_on & step) orc_jet_src(p, s->Src, s->ev[step], tau);
    orc_ideal_update_ev(p, s->ev[3 - step], s->ev[1], s->Src, tau, step, s->iters);
}

/* CLVisc.IS_stepUpdate(step), visc.py:287-314 */
void orc_is_stage(const orc_params *p, orc_state *s, int step, real tau) {
    int bulk = p->bulk_on && s->bulk[0];
    orc_is_christoffel(p, s->IS_src, s->pi[step], s->ev[step], tau, step);
    if (bulk && step == 1) memset(s->bulk_src, 0, sizeof(real) * ncells(p));
    real *uds[3] = {s->udx, s->udy, s->udz};
    for (int a = 0; a < 3; a++)
        orc_is_along(p, a, s->IS_src, uds[a], s->pi[step], s->ev[step],
                     bulk ? s->bulk_src : NULL, bulk ? s->bulk[step] : NULL, tau);
    orc_update_pimn(p, s->pi[3 - step], s->goodcell, s->pi[1], s->pi[step], s->ev[1], s->ev[2],
                    s->udiff, s->udx, s->udy, s->udz, s->IS_src,
                    bulk ? s->bulk[3 - step] : NULL, bulk ? s->bulk[1] : NULL,
                    bulk ? s->bulk[step] : NULL, bulk ? s->bulk_src : NULL, tau, step);
}

/* CLVisc.visc_stepUpdate(step), visc.py:232-283 */
void orc_visc_stage(const orc_params *p, orc_state *s, int step, real tau) {
    int bulk = p->bulk_on && s->bulk[0];
    orc_visc_christoffel(p, s->Src, s->ev[step], s->pi[step], bulk ? s->bulk[step] : NULL, tau, step);
    for (int a = 0; a < 3; a++)
        orc_visc_along(p, a, s->Src, s->ev[step], s->pi[step], bulk ? s->bulk[step] : NULL, tau);
    if (p->jet_on & step) orc_jet_src(p, s->Src, s->ev[step], tau);
    orc_visc_update_ev(p, s->ev[3 - step], s->ev[1], s->pi[0], s->pi[3 - step],
                       bulk ? s->bulk[0] : NULL, bulk ? s->bulk[3 - step] : NULL,
                       s->Src, tau, step, s->iters);
}

/* tau after update_time(loop): cfg.real(TAU0 + (loop+1)*DT), ideal.py:317-320 */
static real tau_of(const orc_params *p, int64_t loop_plus_one) {
    return (real)(p->tau0_cfg + (double)loop_plus_one * p->dt_cfg);
}

/* nsteps of CLIdeal.evolve's step body (ideal.py:341-344) starting at loop index loop0 */
void orc_ideal_steps(const orc_params *p, orc_state *s, int64_t loop0, int64_t nsteps) {
    for (int64_t n = loop0; n < loop0 + nsteps; n++) {
        real tau = tau_of(p, n);
        orc_ideal_stage(p, s, 1, tau);
        tau = tau_of(p, n + 1);
        orc_ideal_stage(p, s, 2, tau);
    }
}

/* nsteps of CLVisc.evolve's step body (visc.py:527-544) starting at loop index loop0 */
void orc_visc_steps(const orc_params *p, orc_state *s, int64_t loop0, int64_t nsteps) {
    int64_t n4 = 4 * ncells(p), n10 = 10 * ncells(p);
    for (int64_t n = loop0; n < loop0 + nsteps; n++) {
        memcpy(s->pi[0], s->pi[1], sizeof(real) * n10);
        memcpy(s->ev[0], s->ev[1], sizeof(real) * n4);
        if (p->bulk_on && s->bulk[0]) memcpy(s->bulk[0], s->bulk[1], sizeof(real) * ncells(p));
        real tau = tau_of(p, n);
        orc_is_stage(p, s, 1, tau);
        orc_visc_stage(p, s, 1, tau);
        tau = tau_of(p, n + 1);
        orc_is_stage(p, s, 2, tau);
        orc_visc_stage(p, s, 2, tau);
        orc_get_udiff(p, s->udiff, s->ev[0], s->ev[1]);
    }
}

int orc_real_size(void) { return (int)sizeof(real); }

/* ------------------------------------------------------------------------- */
/* Gaussian smearing of partons, kernel_gaussian_smearing_new.cl:11-120      */
/* ------------------------------------------------------------------------- */
/* p4x4: np rows (E, px, py, pz, t, x, y, z); one_o_2sigr2 and w (= w1 w2) as the device
 * rounds them; ev: AoS rows of 4.  Per parton: eta_s = (log max(t+z, acu) - log max(t-z,
 * acu))/2, m_T = sqrt(max(E^2 - pz^2, px^2 + py^2)), Y = (log max(E+pz, acu) -
 * log max(E-pz, acu))/2; per cell the partons are summed in list order; then KFACTOR and
 * the fixed-point root solve of helper.h:112-125. */
void orc_smear(const orc_params *p, const real *p4x4, int64_t np_, real one_o_2sigr2, real w, real kfactor,
               real *ev) {
    int nx = p->nx, ny = p->ny, nz = p->nz;
    real *pre = (real *)malloc(sizeof(real) * 7 * (np_ > 0 ? np_ : 1));
    for (int64_t m = 0; m < np_; m++) {
        const real *q = p4x4 + 8 * m;
        real E = q[0], px = q[1], py = q[2], pz = q[3], t = q[4], x = q[5], y = q[6], z = q[7];
        real *o = pre + 7 * m;
        o[0] = x; o[1] = y;
#if ORC_DOUBLE
        o[2] = RC(0.5f) * log(r_max(t + z, ACU)) - RC(0.5f) * log(r_max(t - z, ACU));
        o[3] = sqrt(r_max(E * E - pz * pz, px * px + py * py));
        o[4] = RC(0.5f) * (log(r_max(E + pz, ACU)) - log(r_max(E - pz, ACU)));
#else
        o[2] = 0.5f * logf(r_max(t + z, ACU)) - 0.5f * logf(r_max(t - z, ACU));
        o[3] = sqrtf(r_max(E * E - pz * pz, px * px + py * py));
        o[4] = 0.5f * (logf(r_max(E + pz, ACU)) - logf(r_max(E - pz, ACU)));
#endif
        o[5] = px; o[6] = py;
    }
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < (int64_t)nx * ny * nz; I++) {
        int64_t ij = I / nz;
        int k = (int)(I - ij * nz), i = (int)(ij / ny), j = (int)(ij - (ij / ny) * ny);
        real gx = (real)(i - nx / 2) * RC(p->dx), gy = (real)(j - ny / 2) * RC(p->dy), ge = (real)(k - nz / 2) * RC(p->dz);
        real t0 = 0, t1 = 0, t2 = 0, t3 = 0;
        for (int64_t m = 0; m < np_; m++) {
            const real *q = pre + 7 * m;
            real d1 = gx - q[0], d2 = gy - q[1], d3 = ge - q[2];
            real dist = one_o_2sigr2 * (d1 * d1 + d2 * d2 + d3 * d3);
            if (dist < RC(50.0f)) {
#if ORC_DOUBLE
                real delta = w * exp(-dist);
                t0 = t0 + delta * (q[3] * cosh(q[4] - ge));
                t3 = t3 + delta * (q[3] * sinh(q[4] - ge));
#else
                real delta = w * expf(-dist);
                t0 = t0 + delta * (q[3] * coshf(q[4] - ge));
                t3 = t3 + delta * (q[3] * sinhf(q[4] - ge));
#endif
                t1 = t1 + delta * q[5];
                t2 = t2 + delta * q[6];
            }
        }
        t0 = kfactor * t0; t1 = kfactor * t1; t2 = kfactor * t2; t3 = kfactor * t3;
        real K2 = t1 * t1 + t2 * t2 + t3 * t3;
        real K0 = t0 < ACU ? ACU : t0;
        real E1 = K0, E0;
        int it = 0;
        for (;;) {
            E0 = E1;
            E1 = K0 - K2 / (K0 + eos_P(p, E1));
            it++;
            if (it > 100 || (real)fabs((double)(E1 - E0)) / r_max((real)fabs((double)E1), ACU) < ACU) break;
        }
        real *o = ev + 4 * I;
        o[0] = r_max(RC(0.0f), E1);
        o[1] = o[2] = o[3] = RC(0.0f);
    }
    free(pre);
}

/* ------------------------------------------------------------------------- */
/* Vorticity, kernel_vorticity.cl: omega (:38-99), omega_shear (:143-221),     */
/* omega_accT (:225-337), omega_mu (:103-140)                                 */
/* ------------------------------------------------------------------------- */
typedef struct { real ub[4], uk[4], T; } vq;

static vq vort_q(const orc_params *p, const real *ev) {
    vq q;
    real u[4];
    umu4(ev, u);
    real T = eos_T(p, ev[0]);
    static const real gm[4] = {RC(1.0f), RC(-1.0f), RC(-1.0f), RC(-1.0f)};
    for (int c = 0; c < 4; c++) { q.uk[c] = gm[c] * u[c]; q.ub[c] = DIV(q.uk[c], T); }
    q.T = T;
    return q;
}

/* ev_old, ev_new: AoS rows of 4 (d_ev[0], d_ev[1]); out: n rows of 32 in the order
 * omega(6), omega_shear1(16), omega_shear2(4), omega_accT(6).  An axis of one cell has zero
 * partial derivative (the reference reads past the grid there); the Christoffel terms stay. */
void orc_vorticity(const orc_params *p, const real *ev_old, const real *ev_new, real tau, real *out) {
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const real DT = RC(p->dt), DX = RC(p->dx), DY = RC(p->dy), DZ = RC(p->dz);
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < (int64_t)nx * ny * nz; I++) {
        const int k = (int)(I % nz), j = (int)((I / nz) % ny), i = (int)(I / ((int64_t)nz * ny));
        vq qo = vort_q(p, ev_old + 4 * I), qn = vort_q(p, ev_new + 4 * I);
        const int idx[3] = {i, j, k}, dim[3] = {nx, ny, nz};
        const int64_t stride[3] = {(int64_t)ny * nz, nz, 1};
        const real h2[3] = {RC(2.0f) * DX, RC(2.0f) * DY, RC(2.0f) * DZ * tau};
        const real h1[3] = {DX, DY, DZ * tau};
        real dub[3][4], duk[3][4], dT[3];
        for (int a = 0; a < 3; a++) {
            if (dim[a] == 1) {
                for (int c = 0; c < 4; c++) dub[a][c] = duk[a][c] = RC(0.0f);
                dT[a] = RC(0.0f);
            } else if (idx[a] != 0 && idx[a] != dim[a] - 1) {
                vq qp = vort_q(p, ev_new + 4 * (I + stride[a])), qm = vort_q(p, ev_new + 4 * (I - stride[a]));
                for (int c = 0; c < 4; c++) { dub[a][c] = DIV(qp.ub[c] - qm.ub[c], h2[a]); duk[a][c] = DIV(qp.uk[c] - qm.uk[c], h2[a]); }
                dT[a] = DIV(qp.T - qm.T, h2[a]);
            } else if (idx[a] == 0) {
                vq qp = vort_q(p, ev_new + 4 * (I + stride[a]));
                for (int c = 0; c < 4; c++) { dub[a][c] = DIV(qp.ub[c] - qn.ub[c], h1[a]); duk[a][c] = DIV(qp.uk[c] - qn.uk[c], h1[a]); }
                dT[a] = DIV(qp.T - qn.T, h1[a]);
            } else {
                vq qm = vort_q(p, ev_new + 4 * (I - stride[a]));
                for (int c = 0; c < 4; c++) { dub[a][c] = DIV(qn.ub[c] - qm.ub[c], h1[a]); duk[a][c] = DIV(qn.uk[c] - qm.uk[c], h1[a]); }
                dT[a] = DIV(qn.T - qm.T, h1[a]);
            }
        }
        real *o = out + 32 * I;
        real dbt[4];
        for (int c = 0; c < 4; c++) dbt[c] = DIV(qn.ub[c] - qo.ub[c], DT);
        const real hc = RC(0.5f) * RC(0.19733f);
        o[0] = hc * (dub[1][3] - dub[2][2]);
        o[1] = hc * (dub[2][1] - dub[0][3]);
        o[2] = hc * (dub[0][2] - dub[1][1]);
        o[3] = hc * (dbt[3] - dub[2][0] + DIV(qn.ub[3], tau));
        o[4] = hc * (dub[1][0] - dbt[2]);
        o[5] = hc * (dbt[1] - dub[0][0]);
        real dkt[4], dkz[4];
        for (int c = 0; c < 4; c++) dkt[c] = DIV(qn.uk[c] - qo.uk[c], DT);
        const real chr[4] = {DIV(-qn.uk[3], tau), RC(0.0f), RC(0.0f), DIV(-qn.uk[0], tau)};
        for (int c = 0; c < 4; c++) dkz[c] = duk[2][c] + chr[c];
        const real *rows[4] = {dkt, duk[0], duk[1], dkz};
        for (int a = 0; a < 4; a++)
            for (int c = 0; c < 4; c++) o[6 + 4 * a + c] = RC(0.5f) * rows[a][c];
        real u[4], DU[4];
        umu4(ev_new + 4 * I, u);
        for (int c = 0; c < 4; c++) {
            DU[c] = u[0] * dkt[c] + u[1] * duk[0][c] + u[2] * duk[1][c] + u[3] * dkz[c];
            o[22 + c] = RC(0.5f) * DU[c];
        }
        const real T = qn.T;
        const real dTt = DIV(qn.T - qo.T, DT);
        const real gt = DU[0] - DIV(dTt, T), gx = DU[1] - DIV(dT[0], T), gy = DU[2] - DIV(dT[1], T),
                   gz = DU[3] - DIV(dT[2], T);
        const real *kk = qn.uk;
        o[26] = RC(0.5f) * (kk[2] * gz - kk[3] * gy);
        o[27] = RC(0.5f) * (-kk[1] * gz + kk[3] * gx);
        o[28] = RC(0.5f) * (kk[1] * gy - kk[2] * gx);
        o[29] = RC(0.5f) * (-kk[3] * gt + kk[0] * gz);
        o[30] = RC(0.5f) * (kk[2] * gt - kk[0] * gy);
        o[31] = RC(0.5f) * (-kk[1] * gt + kk[0] * gx);
    }
}

/* omega_mu: out n rows of 4 = (T, omega^x, omega^y, omega^eta), zero where e < 0.1 efrz */
void orc_vorticity_mu(const orc_params *p, const real *ev, const real *om32, real efrz, real *out) {
    int64_t n = (int64_t)p->nx * p->ny * p->nz;
    for (int64_t I = 0; I < n; I++) {
        real u[4];
        umu4(ev + 4 * I, u);
        const real *w = om32 + 32 * I;
        real *o = out + 4 * I;
        o[0] = eos_T(p, ev[4 * I]);
        o[1] = -u[0] * w[0] - u[1] * RC(0.0f) - u[2] * w[3] - u[3] * w[4];
        o[2] = -u[0] * w[1] + u[1] * w[3] - u[2] * RC(0.0f) - u[3] * w[5];
        o[3] = -u[0] * w[2] + u[1] * w[4] + u[2] * w[5] - u[3] * RC(0.0f);
        if ((double)ev[4 * I] < 0.1 * (double)efrz) o[0] = o[1] = o[2] = o[3] = RC(0.0f);
    }
}
