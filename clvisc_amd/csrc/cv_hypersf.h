// cv_hypersf.h -- freeze-out hypersurface finder (pyvisc/kernel/kernel_hypersf.cl:237-660).
//
// One thread per 4D cube spanned by (tau_old, tau_new) x 2^3 cells of the skipped grid.
// Cubes whose 16 corners all lie on one side of EFRZ exit at once; the others find the
// intersections of e = EFRZ with the 32 cube edges, the energy-flow direction, the
// convex-hull tetrahedra and their summed normal (dSigma_mu), and append one element
// (dSigma, v, eta_s), its (t, x, y, z) and, for viscous runs, pi^{mu nu} (and the bulk
// pressure when it is evolved) interpolated to the element's centroid.  Appends go through an atomic counter; elements beyond the
// capacity are counted but not stored (the reference has no capacity check,
// pyvisc/ideal.py:113-117).  Same arithmetic contract as the hydro kernels.
#pragma once
#include "cv_device.h"

namespace cv {

struct SfP {
    int nx, ny, nz, nxskip, nyskip, nzskip;
    int gx, gy, gz;
    int64_t n;                 // cells per field plane
    float dx, dy, dz, efrz;
    int64_t cap;
    int viscous;
};

// 4-vector (t, x, y, z) in the precision R of the surface geometry (float: the
// reference's arithmetic; double: CV_FLAG_SF_FP64)
template <typename R> struct V4 { R x, y, z, w; };
template <typename R> struct SfElem { int id[4]; V4<R> center, norm; };

template <typename R> __device__ __forceinline__ V4<R> f4(R a, R b, R c, R d) { return V4<R>{a, b, c, d}; }
template <typename R> __device__ __forceinline__ V4<R> a4(V4<R> a, V4<R> b) { return f4<R>(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
template <typename R> __device__ __forceinline__ V4<R> s4(V4<R> a, V4<R> b) { return f4<R>(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
template <typename R> __device__ __forceinline__ V4<R> m4(R f, V4<R> a) { return f4<R>(f * a.x, f * a.y, f * a.z, f * a.w); }
template <typename R> __device__ __forceinline__ R dt4(V4<R> a, V4<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
template <typename R> __device__ __forceinline__ V4<R> cube_pos(int c) {
    return f4<R>((R)((c >> 3) & 1), (R)(c & 1), (R)((c >> 1) & 1), (R)((c >> 2) & 1));
}
template <typename R> __device__ __forceinline__ R rcpR(R b) { return (R)1 / b; }
template <typename R> __device__ __forceinline__ R acuR();
template <> __device__ __forceinline__ float acuR<float>() { return CV_ACU; }
template <> __device__ __forceinline__ double acuR<double>() { return 1.0e-14; }

template <typename R>
__device__ SfElem<R> sf_construct(int id0, int id1, int id2, int id3, const V4<R> *ints, V4<R> mc) {
    SfElem<R> sf;
    sf.id[0] = id0; sf.id[1] = id1; sf.id[2] = id2; sf.id[3] = id3;
    sf.center = m4<R>((R)0.25, a4(a4(a4(ints[id0], ints[id1]), ints[id2]), ints[id3]));
    V4<R> vo = s4(sf.center, mc);
    V4<R> a = s4(ints[id1], ints[id0]), b = s4(ints[id2], ints[id0]), c = s4(ints[id3], ints[id0]);
    V4<R> nv = f4<R>(a.y * (b.z * c.w - b.w * c.z) + a.z * (b.w * c.y - b.y * c.w) + a.w * (b.y * c.z - b.z * c.y),
                   -(a.x * (b.z * c.w - b.w * c.z) + a.z * (b.w * c.x - b.x * c.w) + a.w * (b.x * c.z - b.z * c.x)),
                   a.x * (b.y * c.w - b.w * c.y) + a.y * (b.w * c.x - b.x * c.w) + a.w * (b.x * c.y - b.y * c.x),
                   -(a.x * (b.y * c.z - b.z * c.y) + a.y * (b.z * c.x - b.x * c.z) + a.z * (b.x * c.y - b.y * c.x)));
    R proj = dt4(vo, nv);
    if (proj < (R)0) nv = f4<R>(-nv.x, -nv.y, -nv.z, -nv.w);
    sf.norm = nv;
    return sf;
}

template <typename R> __device__ __forceinline__ R sf_lcg(int32_t *seed) {
    int32_t prod = (int32_t)((uint32_t)(*seed) * 16807u);
    *seed = (int32_t)((int64_t)prod % 2147483647);
    return (R)(*seed) * rcpR<R>((R)2147483647);
}

template <typename R> __device__ V4<R> sf_mass_center(const V4<R> *ints, int n) {
    V4<R> c = f4<R>(0, 0, 0, 0);
    for (int i = 0; i < n; i++) c = a4(c, ints[i]);
    R r = rcpR<R>((R)n);
    return f4<R>(c.x * r, c.y * r, c.z * r, c.w * r);
}

template <typename R> __device__ void sf_tiny_move(V4<R> *ints, int n, V4<R> *mc, int gid) {
    int32_t seed = (int32_t)((uint32_t)32117 * (uint32_t)(gid + 1));
    for (int i = 0; i < n - 4; i++)
        for (int j = i + 1; j < n - 3; j++)
            for (int k = j + 1; k < n - 2; k++)
                for (int l = k + 1; l < n - 1; l++)
                    for (int m = l + 1; m < n; m++) {
                        SfElem<R> sf = sf_construct(i, j, k, l, ints, *mc);
                        if (dt4(sf.norm, s4(ints[m], sf.center)) == (R)0) {
                            for (int q = 0; q < n; q++) {
                                R a = (sf_lcg<R>(&seed) - (R)0.5f) * (R)1.0E-7f;
                                R b = (sf_lcg<R>(&seed) - (R)0.5f) * (R)1.0E-7f;
                                R c = (sf_lcg<R>(&seed) - (R)0.5f) * (R)1.0E-7f;
                                R d = (sf_lcg<R>(&seed) - (R)0.5f) * (R)1.0E-7f;
                                ints[q] = a4(ints[q], f4<R>(a, b, c, d));
                            }
                            *mc = sf_mass_center(ints, n);
                            break;
                        }
                    }
}

template <typename R> __device__ bool sf_on_hull(const SfElem<R> &sf, const V4<R> *ints, int n) {
    for (int q = 0; q < n; q++) {
        if (q != sf.id[0] && q != sf.id[1] && q != sf.id[2] && q != sf.id[3]) {
            if (dt4(s4(ints[q], sf.center), sf.norm) >= (R)0) return false;
        }
    }
    return true;
}

template <typename R> __device__ V4<R> sf_energy_flow(const R ed[16], R efrz) {
    V4<R> vl = f4<R>(0, 0, 0, 0), vh = f4<R>(0, 0, 0, 0);
    R elsum = 0, ehsum = 0;
    for (int nn = 0; nn < 2; nn++)
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                for (int k = 0; k < 2; k++) {
                    int id = 8 * nn + 4 * k + 2 * j + i;
                    R dek = efrz - ed[id];
                    R adek = fabs(dek);
                    V4<R> w = f4<R>((R)nn * adek, (R)i * adek, (R)j * adek, (R)k * adek);
                    if (dek > (R)0) { elsum += adek; vl = a4(vl, w); }
                    else { ehsum += adek; vh = a4(vh, w); }
                }
    if (fabs(elsum) > acuR<R>()) { R r = rcpR(elsum); vl = f4<R>(vl.x * r, vl.y * r, vl.z * r, vl.w * r); }
    if (fabs(ehsum) > acuR<R>()) { R r = rcpR(ehsum); vh = f4<R>(vh.x * r, vh.y * r, vh.z * r, vh.w * r); }
    return s4(vl, vh);
}

template <typename R> __device__ V4<R> sf_calc_area(V4<R> *ints, V4<R> ef, int n, int gid) {
    V4<R> mc = sf_mass_center(ints, n);
    V4<R> area = f4<R>(0, 0, 0, 0);
    if (n == 4) {
        SfElem<R> sf = sf_construct(0, 1, 2, 3, ints, mc);
        if (dt4(ef, sf.norm) < (R)0) sf.norm = f4<R>(-sf.norm.x, -sf.norm.y, -sf.norm.z, -sf.norm.w);
        area = sf.norm;
    } else if (n > 4) {
        sf_tiny_move(ints, n, &mc, gid);
        for (int i = 0; i < n - 3; i++)
            for (int j = i + 1; j < n - 2; j++)
                for (int k = j + 1; k < n - 1; k++)
                    for (int l = k + 1; l < n; l++) {
                        SfElem<R> sf = sf_construct(i, j, k, l, ints, mc);
                        if (sf_on_hull(sf, ints, n) && dt4(sf.norm, ef) > (R)0) area = a4(area, sf.norm);
                    }
    }
    R r6 = rcpR((R)6);
    return f4<R>(area.x * r6, area.y * r6, area.z * r6, area.w * r6);
}

template <typename R>
__device__ __forceinline__ void sf_between(R el, R er, int cl, int cr, R efrz, V4<R> *all, int *n) {
    R dE1 = efrz - el, dE2 = efrz - er, dE12 = el - er;
    if (dE1 * dE2 < (R)0) {
        R ratio = fabs(dE1) * rcpR(fabs(dE12));
        all[*n] = a4(m4(ratio, cube_pos<R>(cr)), m4((R)1 - ratio, cube_pos<R>(cl)));
        *n += 1;
    }
}

template <typename R> __device__ int sf_all_intersections(const R ed[16], R efrz, V4<R> *all) {
    int n = 0;
    for (int s = 0; s < 16; s += 4) {
        sf_between(ed[s + 0], ed[s + 1], s + 0, s + 1, efrz, all, &n);
        sf_between(ed[s + 0], ed[s + 2], s + 0, s + 2, efrz, all, &n);
        sf_between(ed[s + 2], ed[s + 3], s + 2, s + 3, efrz, all, &n);
        sf_between(ed[s + 1], ed[s + 3], s + 1, s + 3, efrz, all, &n);
    }
    for (int s = 0; s < 16; s += 8) {
        sf_between(ed[s + 0], ed[s + 4], s + 0, s + 4, efrz, all, &n);
        sf_between(ed[s + 1], ed[s + 5], s + 1, s + 5, efrz, all, &n);
        sf_between(ed[s + 2], ed[s + 6], s + 2, s + 6, efrz, all, &n);
        sf_between(ed[s + 3], ed[s + 7], s + 3, s + 7, efrz, all, &n);
    }
    for (int s = 0; s < 8; s++) sf_between(ed[s], ed[s + 8], s, s + 8, efrz, all, &n);
    return n;
}

template <typename R> __device__ __forceinline__ R cosh_(R x);
template <typename R> __device__ __forceinline__ R sinh_(R x);
template <> __device__ __forceinline__ float cosh_<float>(float x) { return coshf(x); }
template <> __device__ __forceinline__ float sinh_<float>(float x) { return sinhf(x); }
template <> __device__ __forceinline__ double cosh_<double>(double x) { return cosh(x); }
template <> __device__ __forceinline__ double sinh_<double>(double x) { return sinh(x); }

// Cubes (i, j, k) of the thinned grid, k fastest, in two passes: k_hypersf_mark lists the
// cubes that cross EFRZ (one thread per cube), k_hypersf builds their elements (one thread
// per listed cube, so that the wavefronts of the second pass are full).  ev_old/ev_new,
// pi_old/pi_new, bk_old/bk_new: SoA planes of n cells (pi only when viscous, bk only
// with bulk viscosity).  R is the precision of everything computed from the corner
// values; the stored element is rounded to float.

// Corner cells and energy densities of cube t of the skipped grid; false if the cube lies
// outside the grid or all its 16 corners are on one side of EFRZ (the early exit of
// get_hypersf, kernel_hypersf.cl:439-455).
template <typename R>
__device__ __forceinline__ bool sf_cube_corners(const SfP &p, int64_t t, const float *__restrict__ ev_old,
                                                const float *__restrict__ ev_new, int64_t cid[16], R ed[16]) {
    const int k = (int)(t % p.gz);
    const int j = (int)((t / p.gz) % p.gy);
    const int i = (int)(t / ((int64_t)p.gz * p.gy));
    if (!((i + 1) * p.nxskip < p.nx && (j + 1) * p.nyskip < p.ny && (k + 1) * p.nzskip < p.nz)) return false;
    const R EFRZ = (R)p.efrz;
#pragma unroll
    for (int c = 0; c < 16; c++) {
        int ii = (i + (c & 1)) * p.nxskip, jj = (j + ((c >> 1) & 1)) * p.nyskip, kk = (k + ((c >> 2) & 1)) * p.nzskip;
        cid[c] = ((int64_t)ii * p.ny + jj) * p.nz + kk;
        ed[c] = (R)(c < 8 ? ev_old : ev_new)[cid[c]];
    }
    return !((ed[0] - EFRZ) * (ed[15] - EFRZ) > 0 && (ed[1] - EFRZ) * (ed[14] - EFRZ) > 0 &&
             (ed[2] - EFRZ) * (ed[13] - EFRZ) > 0 && (ed[3] - EFRZ) * (ed[12] - EFRZ) > 0 &&
             (ed[4] - EFRZ) * (ed[11] - EFRZ) > 0 && (ed[5] - EFRZ) * (ed[10] - EFRZ) > 0 &&
             (ed[6] - EFRZ) * (ed[9] - EFRZ) > 0 && (ed[7] - EFRZ) * (ed[8] - EFRZ) > 0);
}

// Pass 1: the cubes that cross EFRZ, listed (one atomic per wavefront, slots by ballot).
template <typename R>
__global__ __launch_bounds__(256) void k_hypersf_mark(SfP p, const float *__restrict__ ev_old,
                                                      const float *__restrict__ ev_new, uint32_t *list,
                                                      unsigned *nlist) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)p.gx * p.gy * p.gz;
    bool cross = false;
    if (t < total) {
        int64_t cid[16];
        R ed[16];
        cross = sf_cube_corners<R>(p, t, ev_old, ev_new, cid, ed);
    }
    const uint64_t mask = __ballot(cross);
    if (mask == 0ull) return;                          // uniform over the wavefront
    const int lane = (int)__lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    unsigned base = 0u;
    if (lane == leader) base = atomicAdd(nlist, (unsigned)__popcll(mask));
    base = __shfl(base, leader, 64);
    if (cross) list[base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull))] = (uint32_t)t;
}

// Pass 2: one element per listed cube (grid-stride over the list).
template <typename R>
__device__ void sf_cube(const SfP &p, int64_t t, const float *__restrict__ ev_old, const float *__restrict__ ev_new,
                        const float *__restrict__ pi_old, const float *__restrict__ pi_new,
                        const float *__restrict__ bk_old, const float *__restrict__ bk_new, float tau_old_f,
                        float tau_new_f, unsigned long long *count, float *sf, float *txyz, float *pisf, float *bksf,
                        const float *__restrict__ om_old, const float *__restrict__ om_new, float *omsf);

template <typename R>
__global__ __launch_bounds__(128) void k_hypersf(SfP p, const uint32_t *__restrict__ list,
                                                 const unsigned *__restrict__ nlist, const float *__restrict__ ev_old,
                                                 const float *__restrict__ ev_new,
                                                 const float *__restrict__ pi_old,
                                                 const float *__restrict__ pi_new,
                                                 const float *__restrict__ bk_old,
                                                 const float *__restrict__ bk_new, float tau_old_f, float tau_new_f,
                                                 unsigned long long *count, float *sf, float *txyz, float *pisf,
                                                 float *bksf, const float *__restrict__ om_old = nullptr,
                                                 const float *__restrict__ om_new = nullptr, float *omsf = nullptr) {
    const unsigned m = *nlist;
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < m; q += gridDim.x * blockDim.x)
        sf_cube<R>(p, (int64_t)list[q], ev_old, ev_new, pi_old, pi_new, bk_old, bk_new, tau_old_f, tau_new_f, count,
                   sf, txyz, pisf, bksf, om_old, om_new, omsf);
}

template <typename R>
__device__ void sf_cube(const SfP &p, int64_t t, const float *__restrict__ ev_old, const float *__restrict__ ev_new,
                        const float *__restrict__ pi_old, const float *__restrict__ pi_new,
                        const float *__restrict__ bk_old, const float *__restrict__ bk_new, float tau_old_f,
                        float tau_new_f, unsigned long long *count, float *sf, float *txyz, float *pisf, float *bksf,
                        const float *__restrict__ om_old, const float *__restrict__ om_new, float *omsf) {
    const int k = (int)(t % p.gz);
    const int j = (int)((t / p.gz) % p.gy);
    const int i = (int)(t / ((int64_t)p.gz * p.gy));
    int64_t cid[16];
    R ed[16];
    if (!sf_cube_corners<R>(p, t, ev_old, ev_new, cid, ed)) return;
    const R EFRZ = (R)p.efrz;
    V4<R> all[32];
    int nint = sf_all_intersections<R>(ed, EFRZ, all);
    V4<R> ef = sf_energy_flow<R>(ed, EFRZ);
    V4<R> mc = sf_mass_center<R>(all, nint);
    V4<R> dS = sf_calc_area<R>(all, ef, nint, (int)((uint32_t)i * (uint32_t)j * (uint32_t)k));
    const R tau_old = (R)tau_old_f, tau_new = (R)tau_new_f;
    R dtd = tau_new - tau_old;
    R dxd = (R)p.nxskip * (R)p.dx, dyd = (R)p.nyskip * (R)p.dy, dzd = (R)p.nzskip * (R)p.dz;
    R tau = tau_old + mc.x * dtd;
    R x = (R)(-p.nx / 2) * (R)p.dx + (R)i * dxd + mc.y * dxd;
    R y = (R)(-p.ny / 2) * (R)p.dy + (R)j * dyd + mc.z * dyd;
    R eta = (R)(-p.nz / 2) * (R)p.dz + (R)k * dzd + mc.w * dzd;
    R w[16];
    {
        R a0 = (R)1 - mc.x, a1 = (R)1 - mc.y, a2 = (R)1 - mc.z, a3 = (R)1 - mc.w;
#pragma unroll
        for (int c = 0; c < 16; c++) {
            R tt = (c & 8) ? mc.x : a0, xx = (c & 1) ? mc.y : a1, yy = (c & 2) ? mc.z : a2, zz = (c & 4) ? mc.w : a3;
            w[c] = tt * xx * yy * zz;
        }
    }
    V4<R> evm = f4<R>(0, 0, 0, 0);
#pragma unroll
    for (int c = 0; c < 16; c++) {
        const float *src = c < 8 ? ev_old : ev_new;
        V4<R> e = f4<R>((R)src[cid[c]], (R)src[p.n + cid[c]], (R)src[2 * p.n + cid[c]], (R)src[3 * p.n + cid[c]]);
        evm = a4(evm, m4(w[c], e));
    }
    if (p.viscous && isnan(mc.x)) return;
    unsigned long long id = atomicAdd(count, 1ull);
    if ((int64_t)id >= p.cap) return;
    float *o = sf + 8 * id;
    o[0] = (float)(tau * dxd * dyd * dzd * dS.x);
    o[1] = (float)(-tau * dtd * dyd * dzd * dS.y);
    o[2] = (float)(-tau * dtd * dxd * dzd * dS.z);
    o[3] = (float)(-dtd * dxd * dyd * dS.w);
    o[4] = (float)evm.y; o[5] = (float)evm.z; o[6] = (float)evm.w; o[7] = (float)eta;
    float *tx = txyz + 4 * id;
    tx[0] = (float)(tau * cosh_<R>(eta)); tx[1] = (float)x; tx[2] = (float)y; tx[3] = (float)(tau * sinh_<R>(eta));
    if (p.viscous) {
        for (int mn = 0; mn < 10; mn++) {
            R acc = 0;
#pragma unroll
            for (int c = 0; c < 16; c++) acc = acc + w[c] * (R)(c < 8 ? pi_old : pi_new)[(int64_t)mn * p.n + cid[c]];
            pisf[10 * id + mn] = (float)acc;
        }
        if (bk_old) {
            R acc = 0;
#pragma unroll
            for (int c = 0; c < 16; c++) acc = acc + w[c] * (R)(c < 8 ? bk_old : bk_new)[cid[c]];
            bksf[id] = (float)acc;
        }
    }
    if (om_old) {      // vorticity tensors at the centroid (visc_hypersf, kernel_hypersf.cl:621-650)
        for (int m = 0; m < 32; m++) {
            R acc = 0;
#pragma unroll
            for (int c = 0; c < 16; c++) acc = acc + w[c] * (R)(c < 8 ? om_old : om_new)[(int64_t)m * p.n + cid[c]];
            omsf[32 * id + m] = (float)acc;
        }
    }
}

}  // namespace cv
