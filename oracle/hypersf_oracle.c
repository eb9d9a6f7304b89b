/*
 * hypersf_oracle.c -- CPU oracle of the freeze-out hypersurface finder.
 *
 * TEST INFRASTRUCTURE ONLY (see clvisc_oracle.c).  Restates
 * pyvisc/kernel/kernel_hypersf.cl:
 *   construct_hypersf :91-115   rand / tiny_move_if_coplanar :117-149
 *   is_on_convex_hull :153-167  contribution_from / energy_flow :171-216
 *   get_mass_center   :219-225  calc_area :230-265
 *   ints_between / get_all_intersections :268-314
 *   centroid_ev / centroid_intp :327-380
 *   get_hypersf :405-496 (ideal)  visc_hypersf :501-660 (viscous, pi on the surface)
 * under the same arithmetic contract as clvisc_oracle.c (a/b = a * RN(1/b), no FMA).
 * cosh/sinh of the sf_txyz output come from libm (the reference uses OpenCL's).
 * Elements are appended in loop order (i, j, k with k fastest); the reference appends
 * with atomic_inc, so its order is not defined.  Vorticity on the surface is not
 * part of this path.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

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
static real rcp_(real b) { return RC(1.0f) / b; }
#define DIV(a, b) ((a) * rcp_(b))

typedef struct { real s[4]; } r4;

static r4 mk(real a, real b, real c, real d) { r4 r = {{a, b, c, d}}; return r; }
static r4 add4(r4 a, r4 b) { return mk(a.s[0] + b.s[0], a.s[1] + b.s[1], a.s[2] + b.s[2], a.s[3] + b.s[3]); }
static r4 sub4(r4 a, r4 b) { return mk(a.s[0] - b.s[0], a.s[1] - b.s[1], a.s[2] - b.s[2], a.s[3] - b.s[3]); }
static r4 scl4(real f, r4 a) { return mk(f * a.s[0], f * a.s[1], f * a.s[2], f * a.s[3]); }
static real dot4(r4 a, r4 b) { return a.s[0] * b.s[0] + a.s[1] * b.s[1] + a.s[2] * b.s[2] + a.s[3] * b.s[3]; }

/* corners of the 4D cube: index 8 t + 4 z + 2 y + x, position (t, x, y, z) */
static r4 cube_pos(int c) {
    return mk((real)((c >> 3) & 1), (real)(c & 1), (real)((c >> 1) & 1), (real)((c >> 2) & 1));
}

typedef struct { int id[4]; r4 center, norm; } sf_t;

static sf_t construct(int id0, int id1, int id2, int id3, const r4 *ints, r4 mc) {
    sf_t sf;
    sf.id[0] = id0; sf.id[1] = id1; sf.id[2] = id2; sf.id[3] = id3;
    sf.center = scl4(RC(0.25f), add4(add4(add4(ints[id0], ints[id1]), ints[id2]), ints[id3]));
    r4 vo = sub4(sf.center, mc);
    r4 a = sub4(ints[id1], ints[id0]), b = sub4(ints[id2], ints[id0]), c = sub4(ints[id3], ints[id0]);
    const real *A = a.s, *B = b.s, *Cc = c.s;
    r4 nv = mk(A[1] * (B[2] * Cc[3] - B[3] * Cc[2]) + A[2] * (B[3] * Cc[1] - B[1] * Cc[3]) + A[3] * (B[1] * Cc[2] - B[2] * Cc[1]),
               -(A[0] * (B[2] * Cc[3] - B[3] * Cc[2]) + A[2] * (B[3] * Cc[0] - B[0] * Cc[3]) + A[3] * (B[0] * Cc[2] - B[2] * Cc[0])),
               A[0] * (B[1] * Cc[3] - B[3] * Cc[1]) + A[1] * (B[3] * Cc[0] - B[0] * Cc[3]) + A[3] * (B[0] * Cc[1] - B[1] * Cc[0]),
               -(A[0] * (B[1] * Cc[2] - B[2] * Cc[1]) + A[1] * (B[2] * Cc[0] - B[0] * Cc[2]) + A[2] * (B[0] * Cc[1] - B[1] * Cc[0])));
    real proj = dot4(vo, nv);
    if (proj < RC(0.0f)) nv = mk(-nv.s[0], -nv.s[1], -nv.s[2], -nv.s[3]);
    sf.norm = nv;
    return sf;
}

/* Park-Miller step with the reference's 32-bit int product (wraps) */
static real lcg(int32_t *seed) {
    int32_t prod = (int32_t)((uint32_t)(*seed) * 16807u);
    *seed = (int32_t)((int64_t)prod % 2147483647);
    return DIV((real)(*seed), (real)2147483647);
}

static r4 mass_center(const r4 *ints, int n) {
    r4 c = mk(0, 0, 0, 0);
    for (int i = 0; i < n; i++) c = add4(c, ints[i]);
    real r = rcp_((real)n);
    return mk(c.s[0] * r, c.s[1] * r, c.s[2] * r, c.s[3] * r);
}

static void tiny_move(r4 *ints, int n, r4 *mc, int gid) {
    int32_t seed = (int32_t)((uint32_t)32117 * (uint32_t)(gid + 1));
    for (int i = 0; i < n - 4; i++)
        for (int j = i + 1; j < n - 3; j++)
            for (int k = j + 1; k < n - 2; k++)
                for (int l = k + 1; l < n - 1; l++)
                    for (int m = l + 1; m < n; m++) {
                        sf_t sf = construct(i, j, k, l, ints, *mc);
                        if (dot4(sf.norm, sub4(ints[m], sf.center)) == RC(0.0f)) {
                            for (int q = 0; q < n; q++) {
                                real a = (lcg(&seed) - RC(0.5f)) * RC(1.0E-7f);
                                real b = (lcg(&seed) - RC(0.5f)) * RC(1.0E-7f);
                                real c = (lcg(&seed) - RC(0.5f)) * RC(1.0E-7f);
                                real d = (lcg(&seed) - RC(0.5f)) * RC(1.0E-7f);
                                ints[q] = add4(ints[q], mk(a, b, c, d));
                            }
                            *mc = mass_center(ints, n);
                            break;
                        }
                    }
}

static int on_hull(sf_t sf, const r4 *ints, int n) {
    for (int q = 0; q < n; q++) {
        if (q != sf.id[0] && q != sf.id[1] && q != sf.id[2] && q != sf.id[3]) {
            if (dot4(sub4(ints[q], sf.center), sf.norm) >= RC(0.0f)) return 0;
        }
    }
    return 1;
}

static r4 energy_flow(const real ed[16], real efrz) {
    r4 vl = mk(0, 0, 0, 0), vh = mk(0, 0, 0, 0);
    real elsum = RC(0.0f), ehsum = RC(0.0f);
    for (int n = 0; n < 2; n++)
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                for (int k = 0; k < 2; k++) {
                    int id = 8 * n + 4 * k + 2 * j + i;
                    real dek = efrz - ed[id];
                    real adek = (real)fabs((double)dek);
                    r4 w = mk((real)n * adek, (real)i * adek, (real)j * adek, (real)k * adek);
                    if (dek > RC(0.0f)) { elsum += adek; vl = add4(vl, w); }
                    else { ehsum += adek; vh = add4(vh, w); }
                }
    if ((real)fabs((double)elsum) > ACU) { real r = rcp_(elsum); vl = mk(vl.s[0] * r, vl.s[1] * r, vl.s[2] * r, vl.s[3] * r); }
    if ((real)fabs((double)ehsum) > ACU) { real r = rcp_(ehsum); vh = mk(vh.s[0] * r, vh.s[1] * r, vh.s[2] * r, vh.s[3] * r); }
    return sub4(vl, vh);
}

static r4 calc_area(r4 *ints, r4 ef, int n, int gid) {
    r4 mc = mass_center(ints, n);
    r4 area = mk(0, 0, 0, 0);
    if (n == 4) {
        sf_t sf = construct(0, 1, 2, 3, ints, mc);
        if (dot4(ef, sf.norm) < RC(0.0f)) sf.norm = mk(-sf.norm.s[0], -sf.norm.s[1], -sf.norm.s[2], -sf.norm.s[3]);
        area = sf.norm;
    } else if (n > 4) {
        tiny_move(ints, n, &mc, gid);
        for (int i = 0; i < n - 3; i++)
            for (int j = i + 1; j < n - 2; j++)
                for (int k = j + 1; k < n - 1; k++)
                    for (int l = k + 1; l < n; l++) {
                        sf_t sf = construct(i, j, k, l, ints, mc);
                        if (on_hull(sf, ints, n) && dot4(sf.norm, ef) > RC(0.0f)) area = add4(area, sf.norm);
                    }
    }
    real r6 = rcp_(RC(6.0f));
    return mk(area.s[0] * r6, area.s[1] * r6, area.s[2] * r6, area.s[3] * r6);
}

static void ints_between(real el, real er, r4 pl, r4 pr, real efrz, r4 *all, int *n) {
    real dE1 = efrz - el, dE2 = efrz - er, dE12 = el - er;
    if (dE1 * dE2 < RC(0.0f)) {
        real ratio = DIV((real)fabs((double)dE1), (real)fabs((double)dE12));
        all[*n] = add4(scl4(ratio, pr), scl4(RC(1.0f) - ratio, pl));
        *n += 1;
    }
}

static int all_intersections(const real ed[16], real efrz, r4 *all) {
    int n = 0;
    for (int s = 0; s < 16; s += 4) {
        ints_between(ed[s + 0], ed[s + 1], cube_pos(s + 0), cube_pos(s + 1), efrz, all, &n);
        ints_between(ed[s + 0], ed[s + 2], cube_pos(s + 0), cube_pos(s + 2), efrz, all, &n);
        ints_between(ed[s + 2], ed[s + 3], cube_pos(s + 2), cube_pos(s + 3), efrz, all, &n);
        ints_between(ed[s + 1], ed[s + 3], cube_pos(s + 1), cube_pos(s + 3), efrz, all, &n);
    }
    for (int s = 0; s < 16; s += 8) {
        ints_between(ed[s + 0], ed[s + 4], cube_pos(s + 0), cube_pos(s + 4), efrz, all, &n);
        ints_between(ed[s + 1], ed[s + 5], cube_pos(s + 1), cube_pos(s + 5), efrz, all, &n);
        ints_between(ed[s + 2], ed[s + 6], cube_pos(s + 2), cube_pos(s + 6), efrz, all, &n);
        ints_between(ed[s + 3], ed[s + 7], cube_pos(s + 3), cube_pos(s + 7), efrz, all, &n);
    }
    for (int s = 0; s < 8; s++)
        ints_between(ed[s], ed[s + 8], cube_pos(s), cube_pos(s + 8), efrz, all, &n);
    return n;
}

/* 16-corner weights of the 4D linear interpolation at mc, in the reference's order */
static void corner_weights(r4 mc, real w[16]) {
    real a0 = RC(1.0f) - mc.s[0], a1 = RC(1.0f) - mc.s[1], a2 = RC(1.0f) - mc.s[2], a3 = RC(1.0f) - mc.s[3];
    real b0 = mc.s[0], b1 = mc.s[1], b2 = mc.s[2], b3 = mc.s[3];
    for (int c = 0; c < 16; c++) {
        real t = (c & 8) ? b0 : a0, x = (c & 1) ? b1 : a1, y = (c & 2) ? b2 : a2, z = (c & 4) ? b3 : a3;
        w[c] = t * x * y * z;
    }
}

typedef struct {
    int32_t nx, ny, nz, nxskip, nyskip, nzskip;
    float dx, dy, dz, efrz;
} orc_sf_params;

/* one pass of get_hypersf / visc_hypersf over the cubes between ev_old (tau_old) and
 * ev_new (tau_new); AoS inputs; appends to sf (8/elem), txyz (4/elem), pisf (10/elem,
 * viscous only) and bksf (1/elem, bulk pressure at the centroid; viscous with bk_old
 * only -- the reference evolves no bulk pressure) starting at *count; returns the
 * number appended (bounded by cap). */
int64_t orc_hypersf(const orc_sf_params *p, const real *ev_old, const real *ev_new,
                    const real *pi_old, const real *pi_new, const real *bk_old, const real *bk_new,
                    real tau_old, real tau_new, real *sf, real *txyz, real *pisf, real *bksf,
                    int64_t *count, int64_t cap, int viscous) {
    const int NX = p->nx, NY = p->ny, NZ = p->nz;
    const int gx = (NX - 1) / p->nxskip + 1, gy = (NY - 1) / p->nyskip + 1, gz = (NZ - 1) / p->nzskip + 1;
    const real EFRZ = RC(p->efrz);
    int64_t added = 0;
    for (int i = 0; i < gx; i++)
        for (int j = 0; j < gy; j++)
            for (int k = 0; k < gz; k++) {
                if (!((i + 1) * p->nxskip < NX && (j + 1) * p->nyskip < NY && (k + 1) * p->nzskip < NZ)) continue;
                int64_t cid[16];
                for (int c = 0; c < 16; c++) {
                    int ii = (i + (c & 1)) * p->nxskip, jj = (j + ((c >> 1) & 1)) * p->nyskip,
                        kk = (k + ((c >> 2) & 1)) * p->nzskip;
                    cid[c] = ((int64_t)ii * NY + jj) * NZ + kk;
                }
                real ed[16];
                r4 evc[16];
                for (int c = 0; c < 16; c++) {
                    const real *src = (c < 8 ? ev_old : ev_new) + 4 * cid[c];
                    evc[c] = mk(src[0], src[1], src[2], src[3]);
                    ed[c] = src[0];
                }
                int surf = 1;
                if ((ed[0] - EFRZ) * (ed[15] - EFRZ) > 0 && (ed[1] - EFRZ) * (ed[14] - EFRZ) > 0 &&
                    (ed[2] - EFRZ) * (ed[13] - EFRZ) > 0 && (ed[3] - EFRZ) * (ed[12] - EFRZ) > 0 &&
                    (ed[4] - EFRZ) * (ed[11] - EFRZ) > 0 && (ed[5] - EFRZ) * (ed[10] - EFRZ) > 0 &&
                    (ed[6] - EFRZ) * (ed[9] - EFRZ) > 0 && (ed[7] - EFRZ) * (ed[8] - EFRZ) > 0)
                    surf = 0;
                if (!surf) continue;
                r4 all[32];
                int nint = all_intersections(ed, EFRZ, all);
                r4 ef = energy_flow(ed, EFRZ);
                r4 mc = mass_center(all, nint);
                r4 dS = calc_area(all, ef, nint, (int)((uint32_t)i * (uint32_t)j * (uint32_t)k));
                real dtd = tau_new - tau_old;
                real dxd = (real)p->nxskip * RC(p->dx), dyd = (real)p->nyskip * RC(p->dy), dzd = (real)p->nzskip * RC(p->dz);
                real tau = tau_old + mc.s[0] * dtd;
                real x = (real)(-NX / 2) * RC(p->dx) + (real)i * dxd + mc.s[1] * dxd;
                real y = (real)(-NY / 2) * RC(p->dy) + (real)j * dyd + mc.s[2] * dyd;
                real eta = (real)(-NZ / 2) * RC(p->dz) + (real)k * dzd + mc.s[3] * dzd;
                real w[16];
                corner_weights(mc, w);
                r4 evm = mk(0, 0, 0, 0);
                for (int c = 0; c < 16; c++) evm = add4(evm, scl4(w[c], evc[c]));
                if (viscous && isnan((double)mc.s[0])) continue;
                int64_t id = *count;
                *count += 1;
                if (id >= cap) continue;
                real *o = sf + 8 * id;
                o[0] = tau * dxd * dyd * dzd * dS.s[0];
                o[1] = -tau * dtd * dyd * dzd * dS.s[1];
                o[2] = -tau * dtd * dxd * dzd * dS.s[2];
                o[3] = -dtd * dxd * dyd * dS.s[3];
                o[4] = evm.s[1]; o[5] = evm.s[2]; o[6] = evm.s[3]; o[7] = eta;
                real *t = txyz + 4 * id;
                t[0] = tau * (real)cosh((double)eta); t[1] = x; t[2] = y; t[3] = tau * (real)sinh((double)eta);
                if (viscous && pisf) {
                    for (int mn = 0; mn < 10; mn++) {
                        real acc = RC(0.0f);
                        for (int c = 0; c < 16; c++)
                            acc = acc + w[c] * ((c < 8 ? pi_old : pi_new)[10 * cid[c] + mn]);
                        pisf[10 * id + mn] = acc;
                    }
                    if (bk_old && bksf) {
                        real acc = RC(0.0f);
                        for (int c = 0; c < 16; c++) acc = acc + w[c] * (c < 8 ? bk_old : bk_new)[cid[c]];
                        bksf[id] = acc;
                    }
                }
                added++;
            }
    return added;
}

/* test_hypersf (kernel_hypersf.cl:665-690): corners at tau_old e = 3, at tau_new e = 2 */
void orc_test_hypersf(real efrz, real out[4]) {
    real ed[16];
    for (int i = 0; i < 8; i++) { ed[i] = RC(3.0f); ed[8 + i] = RC(2.0f); }
    r4 all[32];
    int n = all_intersections(ed, efrz, all);
    r4 ef = energy_flow(ed, efrz);
    r4 a = calc_area(all, ef, n, 0);
    for (int c = 0; c < 4; c++) out[c] = a.s[c];
}
