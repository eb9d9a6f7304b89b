/* CPU oracle of the optical Glauber initial condition (TEST INFRASTRUCTURE ONLY).
 *
 * Restates pyvisc/kernel/kernel_glauber.cl: thickness (:10-33, five-point Gauss-Legendre on
 * ten sub-intervals of [0, 5R]), Nw / Nb (:72-81), ed_transverse (:51-54),
 * weight_along_eta (:36-48, etas0 = 0 as :58-66 returns), glauber_ini (:82-102) and
 * num_of_binary_collisions (:104-121), in `real` (float in liborc_f32, double in
 * liborc_f64) with the reference's operation order.  Pinned in tests/test_glauber_oracle.py
 * by scipy quadrature of the Woods-Saxon integral, the nucleus normalisation and the
 * number of participants. */
#include <math.h>
#include <stdint.h>

#ifndef ORC_DOUBLE
#define ORC_DOUBLE 0
#endif
#if ORC_DOUBLE
typedef double real;
#define EXP exp
#define POW pow
#define SQRT sqrt
#define FABS fabs
#else
typedef float real;
#define EXP expf
#define POW powf
#define SQRT sqrtf
#define FABS fabsf
#endif
#define RC(x) ((real)(x))

typedef struct {                 /* the layout of cv_glauber (include/clvisc_amd.h) */
    float num_nucleons, sqrts, ro0, ra, eta, si0, b, edmax, hwn, eta_flat, eta_gw;
    int32_t bjorken;
} orc_glauber;

real orc_glauber_thickness(const orc_glauber *g, real x, real y) {
    static const float zk[5] = {-0.9061798f, -0.5386492f, 0.0f, 0.5386493f, 0.9061798f};
    static const float Ak[5] = {0.2369269f, 0.4786287f, 0.5688889f, 0.4786287f, 0.2369269f};
    real cut = RC(5.0f) * RC(g->ra);
    real t = RC(0.0f);
    for (int j = 0; j != 10; ++j) {
        real a = (real)j * cut / RC(10.0f), b = (real)(j + 1) * cut / RC(10.0f);
        for (int i = 0; i != 5; ++i) {
            real z = (a + b) / RC(2.0f) + (b - a) / RC(2.0f) * RC(zk[i]);
            real r = SQRT(x * x + y * y + z * z);
            real f = RC(g->ro0) / (RC(1.0f) + EXP((r - RC(g->ra)) / RC(g->eta)));
            t = t + RC(Ak[i]) * f * (b - a) / RC(2.0f);
        }
    }
    return RC(2.0f) * t;
}

static real nw(const orc_glauber *g, real ta, real tb) {
    real A = RC(g->num_nucleons), s = RC(g->si0);
    return ta * (RC(1.0f) - POW(RC(1.0f) - s * tb / A, A)) + tb * (RC(1.0f) - POW(RC(1.0f) - s * ta / A, A));
}
static real nb(const orc_glauber *g, real ta, real tb) { return RC(g->si0) * ta * tb; }
static real ed_transverse(const orc_glauber *g, real ta, real tb) {
    return RC(g->hwn) * nw(g, ta, tb) + (RC(1.0f) - RC(g->hwn)) * nb(g, ta, tb);
}
static real weight_eta(const orc_glauber *g, real z) {
    if (g->bjorken) return RC(1.0f);
    real d = FABS(z - RC(0.0f));
    if (d > RC(g->eta_flat))
        return EXP(-POW(d - RC(g->eta_flat), RC(2.0f)) / (RC(2.0f) * RC(g->eta_gw) * RC(g->eta_gw)));
    return RC(1.0f);
}

real orc_glauber_nw(const orc_glauber *g, real ta, real tb) { return nw(g, ta, tb); }

/* ev: nx * ny * nz rows of 4 (AoS, the reference's d_ev1 layout) */
void orc_glauber_ini(const orc_glauber *g, int nx, int ny, int nz, real dx, real dy, real dz, real *ev) {
    real tcent = orc_glauber_thickness(g, RC(0.0f), RC(0.0f));
    real ed_central = ed_transverse(g, tcent, tcent);
    real kfac = RC(g->edmax) / ed_central;
    real b = RC(g->b);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            real x = (real)(i - nx / 2) * dx, y = (real)(j - ny / 2) * dy;
            real ta = orc_glauber_thickness(g, x - RC(0.5f) * b, y);
            real tb = orc_glauber_thickness(g, x + RC(0.5f) * b, y);
            real edxy = kfac * ed_transverse(g, ta, tb);
            for (int k = 0; k < nz; k++) {
                real etas = (real)(k - nz / 2) * dz;
                real *o = ev + 4 * (((int64_t)i * ny + j) * nz + k);
                o[0] = edxy * weight_eta(g, etas);
                o[1] = o[2] = o[3] = RC(0.0f);
            }
        }
}

void orc_glauber_nbinary(const orc_glauber *g, real xlow, real ylow, real dx, real dy, int nx, int ny, real *out) {
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            real x = xlow + (real)i * dx, y = ylow + (real)j * dy;
            real ta = orc_glauber_thickness(g, x - RC(0.5f) * RC(g->b), y);
            real tb = orc_glauber_thickness(g, x + RC(0.5f) * RC(g->b), y);
            out[i * ny + j] = nb(g, ta, tb);
        }
}
