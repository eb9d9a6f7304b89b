/* CPU oracle of the Lambda polarization on the freeze-out surface (TEST INFRASTRUCTURE ONLY).
 *
 * Restates pyvisc/kernel/kernel_polarization.cl (polarization_on_sf :14-148,
 * polarization_shear_on_sf :151-338, polarization_accT_on_sf :341-495) for one momentum at
 * a time in the reference's summation order: terms of 256 consecutive elements reduced by
 * the work-group tree (s = 128, 64, ..., 1), chunks added in order.  `real` is float in
 * liborc_f32, double in liborc_f64.  Pinned in tests/test_polarization_oracle.py by the
 * density's closed form (a static isotropic surface gives the Boltzmann-like integral of
 * feq) and by symmetry (no polarization without vorticity; P odd under p_x -> -p_x for an
 * omega^{tz} field). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#ifndef ORC_DOUBLE
#define ORC_DOUBLE 0
#endif
#if ORC_DOUBLE
typedef double real;
#define EXP exp
#define COSH cosh
#define SINH sinh
#define SQRT sqrt
#else
typedef float real;
#define EXP expf
#define COSH coshf
#define SINH sinhf
#define SQRT sqrtf
#endif
#define RC(x) ((real)(x))
#define BSZ 256

static real dot4(const real *a, const real *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }

/* one element's (density term, 4 polarization terms) */
static void term(int kind, const float *s, const float *w1, const float *w2, real mt, real Y, real px, real py,
                 real tfrz, real beta, real out[5]) {
    real vx = s[4], vy = s[5], vz = s[6], etas = s[7];
    real gi = RC(1.0f) / SQRT(fmax(1.0E-10, (double)(RC(1.0f) - vx * vx - vy * vy - vz * vz)));
    real umu[4] = {RC(1.0f) * gi, vx * gi, vy * gi, vz * gi};
    real umu1[4] = {RC(1.0f) * gi, -vx * gi, -vy * gi, -vz * gi};
    real dsig[4] = {s[0], s[1], s[2], s[3]};
    real mtcosh = mt * COSH(Y - etas), mtsinh = mt * SINH(Y - etas);
    real p1[4] = {mtcosh, -px, -py, -mtsinh};
    real p2[4] = {mtcosh, px, py, mtsinh};
    real pdotu = dot4(umu, p1);
    real tmp = EXP(-pdotu * beta);
    real feq = tmp / (RC(1.0f) + tmp);
    real volum = dot4(dsig, p1);
    real otx, oty, otz, oxy, oxz, oyz, q[4];
    if (kind == 1) {
        real sym[4][4];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) sym[a][b] = a == b ? RC(w1[5 * a]) * RC(2.0f) : RC(w1[4 * b + a]) + RC(w1[4 * a + b]);
        for (int a = 0; a < 4; a++) q[a] = dot4(p2, sym[a]) - pdotu * RC(w2[a]);
        otx = p1[2] * umu1[3] - p1[3] * umu1[2];
        oty = p1[3] * umu1[1] - p1[1] * umu1[3];
        otz = p1[1] * umu1[2] - p1[2] * umu1[1];
        oxy = p1[0] * umu1[3] - p1[3] * umu1[0];
        oxz = p1[2] * umu1[0] - p1[0] * umu1[2];
        oyz = p1[0] * umu1[1] - p1[1] * umu1[0];
    } else {
        otx = w1[0]; oty = w1[1]; otz = w1[2]; oxy = w1[3]; oxz = w1[4]; oyz = w1[5];
        for (int a = 0; a < 4; a++) q[a] = p1[a];
    }
    real pw[4] = {otx * q[1] + oty * q[2] + otz * q[3], -otx * q[0] + oxy * q[2] + oxz * q[3],
                  -oty * q[0] - oxy * q[1] + oyz * q[3], -otz * q[0] - oxz * q[1] - oyz * q[2]};
    real sp[4];
    for (int a = 0; a < 4; a++) {
        if (kind == 0) sp[a] = volum * pw[a] * feq * (RC(1.0f) - feq);
        else if (kind == 1) sp[a] = RC(-0.19733f) * volum * pw[a] * feq * (RC(1.0f) - feq) / (pdotu * tfrz);
        else sp[a] = RC(-0.19733f) * volum * pw[a] * feq * (RC(1.0f) - feq) / tfrz;
    }
    out[0] = volum * feq;
    out[1] = COSH(etas) * sp[0] + SINH(etas) * sp[3];
    out[2] = sp[1];
    out[3] = sp[2];
    out[4] = SINH(etas) * sp[0] + COSH(etas) * sp[3];
}

void orc_polarization(int kind, const float *sf, const float *om1, const float *om2, int64_t nsf, const float *mom,
                      int64_t nmom, real tfrz, real mass, real *pol_lab, real *density, real *pol_lrf) {
    const int w1n = kind == 1 ? 16 : 6;
    const real beta = (real)(1.0 / (double)tfrz);
    #pragma omp parallel for schedule(dynamic)
    for (int64_t g = 0; g < nmom; g++) {
        real mt = mom[4 * g], Y = mom[4 * g + 1], px = mom[4 * g + 2], py = mom[4 * g + 3];
        real acc[5] = {0, 0, 0, 0, 0};
        real (*buf)[5] = malloc(sizeof(real) * 5 * BSZ);
        for (int64_t c0 = 0; c0 < nsf; c0 += BSZ) {
            for (int l = 0; l < BSZ; l++) {
                int64_t i = c0 + l;
                if (i < nsf) term(kind, sf + 8 * i, om1 + (int64_t)w1n * i, kind == 1 ? om2 + 4 * i : NULL, mt, Y, px,
                                  py, tfrz, beta, buf[l]);
                else for (int c = 0; c < 5; c++) buf[l][c] = RC(0.0f);
            }
            for (int st = BSZ >> 1; st > 0; st >>= 1)
                for (int l = 0; l < st; l++)
                    for (int c = 0; c < 5; c++) buf[l][c] += buf[l + st][c];
            for (int c = 0; c < 5; c++) acc[c] += buf[0][c];
        }
        free(buf);
        density[g] = acc[0];
        for (int c = 0; c < 4; c++) pol_lab[4 * g + c] = acc[1 + c] / (RC(4.0f) * mass);
        real L[4];
        for (int c = 0; c < 4; c++) L[c] = acc[1 + c] / (RC(4.0f) * mass * acc[0]);
        real p0 = mt * COSH(Y), pz = mt * SINH(Y);
        real pm[4] = {p0, -px, -py, -pz};
        real Pi0 = dot4(pm, L) / mass;
        real d3 = pm[1] * L[1] + pm[2] * L[2] + pm[3] * L[3];
        pol_lrf[4 * g] = Pi0;
        for (int c = 1; c < 4; c++) pol_lrf[4 * g + c] = L[c] - d3 * pm[c] / (p0 * (p0 + mass));
    }
}
