// clvisc_amd.hip -- C-ABI of the MI355X KT + Israel-Stewart hydro step.
// Declarations and the reference interface each entry point replaces: include/clvisc_amd.h.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include <algorithm>
#include <mutex>

#include "../../include/clvisc_amd.h"
#include "cv_kernels.h"
#include "cv_tiled.h"
#include "cv_ideal_tiled.h"
#include "cv_hypersf.h"
#include "cv_mlp.h"
#include "cv_ic.h"
#include "cv_vort.h"
#include "cv_pol.h"

// learned root-solve seed network (parameters, Adam moments, training set)
struct cv_mlp {
    int device = 0;
    float *P = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr, *loss = nullptr;
    float *part = nullptr, *part_loss = nullptr;
    int64_t part_blocks = 0;
    float *X = nullptr, *T = nullptr;       // training set
    int64_t ndata = 0;
    uint32_t *pk = nullptr;                 // packed BF16 operands of the seed pass (k_mlp_pack_t)
    int64_t adam_t = 0;
    hipStream_t stream = nullptr;
};

using namespace cv;

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                 \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(CV_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

struct cv_handle {
    cv_params prm;
    JetP jet = {};       // jet energy-loss source of the stage kernels (on = 0: none)
    int device;
    int viscous;
    int bulk;
    int64_t n;
    PhysP phys;
    EosP eos;
    float4 *d_eos = nullptr;
    // slots of the reference buffers -> physical device buffers
    float *ev[3] = {nullptr, nullptr, nullptr};
    float *pi[3] = {nullptr, nullptr, nullptr};
    float *bk[3] = {nullptr, nullptr, nullptr};
    float *Src = nullptr, *IS_src = nullptr, *bsrc = nullptr, *udiff = nullptr, *good = nullptr;
    unsigned int *d_max = nullptr;
    double *d_esum = nullptr;     // FP64 energy sums: block partials, then the two totals
    FixEntry *d_fix = nullptr;    // deferred FP64 root solves (CV_FLAG_FP64_ROOT, tiled kernels)
    unsigned *d_fix_ctr = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    bool timed = false;
    int cus = 0;
    int kernel = 2;           // step path: 0 per-cell fused kernel, 1 tiled kernel, 2 tiled x-march
    TileP tile;
    int tile_grid = 0;
    int march_xc = 0;         // x planes per march chunk (0: automatic)
    int xcd = 8;              // XCD-contiguous work mapping of the tiled kernels (MI355X: 8 XCDs; cv_set_xcd_map)
    // learned root-solve seed of the ideal stage (cv_set_root_seed)
    struct cv_mlp *seed_mlp = nullptr;
    bool track_iters = false;
    float *seed_buf = nullptr;  // 10 n floats: T00, M, T03, features (4), outputs (2), seed
                                // (split viscous stage: cons planes 0..3, seed at 9 n)
    bool root_split = false;    // viscous stages in split form (cv_set_root_split)
    int *iters = nullptr;
    // freeze-out surface: fields at the previous surface time, appended elements
    float *sf_ev = nullptr, *sf_pi = nullptr, *sf_bk = nullptr;
    float sf_tau_old = 0.0f;
    bool sf_have_old = false;
    float *sf_out = nullptr, *sf_txyz = nullptr, *sf_pisf = nullptr, *sf_bksf = nullptr;
    // vorticity (cv_vorticity_enable): 32 planes per slot (d_omega[0], d_omega[1] with the
    // shear and accT tensors), and 32 floats per surface element
    float *om[2] = {nullptr, nullptr};
    float *sf_om = nullptr;
    unsigned long long *sf_count = nullptr;
    double *d_slice = nullptr;           // cv_slice_sums partials and sums (first call)
    uint32_t *sf_list = nullptr;         // cubes that cross EFRZ (pass 1 of the search)
    unsigned *sf_nlist = nullptr;
    int64_t sf_list_cap = 0;
    int64_t sf_cap = 0;
};

// The x-marching kernels hold the pressure column of an octave EOS table in LDS: they run
// for row/column tables and for octave tables of at most CV_PTAB_LDS nodes.
static bool march_ok(const cv_handle *h) { return h->eos.kind != 1 || h->eos.nlast + 1 <= CV_PTAB_LDS; }

// Tile of the tiled stage kernel: ty x tz cells of one x-plane with a 2-cell halo in y and
// z (none along an axis of length 1).  Checks the LDS capacities of cv_tiled.h.
static int set_tile(cv_handle *h, int ty, int tz) {
    const cv_params &p = h->prm;
    if (ty < 1 || tz < 1 || ty * tz > CV_TILE_THREADS)
        return fail(CV_ERR_ARG, "tile %dx%d: need 1 <= ty*tz <= %d", ty, tz, CV_TILE_THREADS);
    TileP t;
    t.ty = ty; t.tz = tz;
    t.hy = p.ny > 1 ? 2 : 0;
    t.hz = p.nz > 1 ? 2 : 0;
    t.HY = ty + 2 * t.hy; t.HZ = tz + 2 * t.hz;
    if (t.HY * t.HZ > CV_TILE_MAXC) return fail(CV_ERR_ARG, "tile %dx%d: %d halo cells exceed %d", ty, tz, t.HY * t.HZ, CV_TILE_MAXC);
    if ((t.hy > 0 && (ty + 1) * tz > CV_TILE_MAXF) || (t.hz > 0 && ty * (tz + 1) > CV_TILE_MAXF))
        return fail(CV_ERR_ARG, "tile %dx%d: too many faces for the LDS face buffer", ty, tz);
    t.ntj = (p.ny + ty - 1) / ty;
    t.ntk = (p.nz + tz - 1) / tz;
    t.ntiles = (int64_t)p.nx * t.ntj * t.ntk;
    // march: chunks of xc planes so that about 32 work items per CU exist (3 run at once;
    // short chunks balance the tail, each chunk start re-evaluates one x face)
    const int64_t cols = (int64_t)t.ntj * t.ntk;
    const int64_t target = (int64_t)(h->cus > 0 ? h->cus : 256) * 32;
    int64_t nch = (target + cols - 1) / cols;
    if (nch < 1) nch = 1;
    if (nch > p.nx) nch = p.nx;
    t.xc = (int)((p.nx + nch - 1) / nch);
    if (h->march_xc > 0) t.xc = h->march_xc < p.nx ? h->march_xc : p.nx;
    t.nchunks = (p.nx + t.xc - 1) / t.xc;
    t.xcd = h->xcd;
    // automatic chunks with the XCD-contiguous mapping: one x slab per XCD (an XCD's range
    // is then exactly its slab) whose last chunk is short, so that the last work items an
    // XCD runs are short and the dispatch drains sooner: about 24 work items per CU, i.e.
    // `rows` chunks per slab, the last one xt planes (a twelfth of the slab, 1 to 4) and the
    // others sharing the rest (385 planes, 195 tile columns: 15, 15, 15, 4; at 9-plane
    // chunks the step took 1.2 % longer, with a full-length last chunk 2 % longer).  Where
    // the main chunks would be shorter than twice the tail (few tile columns: the items are
    // short anyway), and for a chunk length set by cv_set_march_chunk, the chunks run along
    // the whole x axis, as before.
    t.slabs = 0; t.cps = 0; t.xt = 0;
    if (h->march_xc == 0 && t.xcd > 1 && p.nx >= 2 * t.xcd) {
        const int maxl = (p.nx + t.xcd - 1) / t.xcd;
        const int xt = std::min(4, std::max(1, maxl / 12));
        const int64_t items = (int64_t)(h->cus > 0 ? h->cus : 256) * 24;
        int64_t rows = (items / std::max<int64_t>(1, cols) + t.xcd / 2) / t.xcd;
        rows = std::max<int64_t>(2, std::min<int64_t>(rows, maxl - xt));   // no empty chunk
        const int nm = (int)rows - 1;
        const int xcm = (maxl - xt + nm - 1) / nm;                          // longest main chunk
        if (maxl - xt >= 2 && (maxl - xt) / nm >= 2 * xt) {
            t.slabs = t.xcd;
            t.xt = xt;
            t.xc = xcm;
            t.cps = nm + 1;
            t.nchunks = t.slabs * t.cps;
        }
    }
    t.nwork = cols * t.nchunks;
    h->tile = t;
    int64_t g = (int64_t)(h->cus > 0 ? h->cus : 256) * 2 * 4;
    h->tile_grid = (int)(t.ntiles < g ? t.ntiles : g);
    return CV_OK;
}

// default tile: least modelled cost per cell update.  Per plane a workgroup runs the halo
// pass (HY x HZ cells), the y and z face passes and the interior pass (x faces, update_pimn,
// root solve) over its 256 threads; each pass costs its VALU work per wave times the
// wavefronts it occupies (work), and the passes a single wavefront repeats while the rest
// of the workgroup waits at the barrier add to the critical path (crit).  Weights are the
// passes' instruction counts per wavefront (halo cell 280, face 640, interior cell 3000).
// Padding of (ny, nz) to whole tiles and a short z run (fewer cells per cache line on the
// x-neighbour loads) scale the result.
static int default_tile(cv_handle *h) {
    const cv_params &p = h->prm;
    // one row of the (y, eta) plane per tile: at most 252 cells, so that the halo pass
    // (ty + 4 cells) and the y-face pass (ty + 1 faces) fit one pass of 256 threads
    if (p.nz == 1) return set_tile(h, p.ny >= 252 ? 252 : p.ny, 1);
    // likewise one eta row when ny == 1 (the halo and the z-face pass then hold tz + 4 cells
    // and tz + 1 faces)
    if (p.ny == 1) return set_tile(h, 1, p.nz >= 252 ? 252 : p.nz);
    auto waves = [](int n, int w) { return (double)((n + w - 1) / w); };
    int best_ty = 0, best_tz = 0;
    double best = 1e30;
    for (int tz = 4; tz <= 64; tz++) {
        for (int ty = 1; ty * tz <= CV_TILE_THREADS; ty++) {
            const int hy = p.ny > 1 ? 2 : 0;
            const int hc = (ty + 2 * hy) * (tz + 4), yf = hy > 0 ? (ty + 1) * tz : 0, zf = ty * (tz + 1);
            if (hc > CV_TILE_MAXC || yf > CV_TILE_MAXF || zf > CV_TILE_MAXF) continue;
            if (ty * tz < 160) continue;
            const double pad =
                (double)((p.ny + ty - 1) / ty * ty) * ((p.nz + tz - 1) / tz * tz) / ((double)p.ny * p.nz);
            const double work = 280.0 * waves(hc, 64) + 640.0 * (waves(yf, 64) + waves(zf, 64)) +
                                3000.0 * waves(ty * tz, 64);
            const double crit = 280.0 * waves(hc, CV_TILE_THREADS) +
                                640.0 * (waves(yf, CV_TILE_THREADS) + waves(zf, CV_TILE_THREADS)) + 3000.0;
            const double score = pad * (work + crit) / (ty * tz) * (1.0 + 0.5 / tz);
            if (score < best) { best = score; best_ty = ty; best_tz = tz; }
        }
    }
    if (!best_ty) return set_tile(h, 8, 32);
    return set_tile(h, best_ty, best_tz);
}

extern "C" const char *cv_last_error(void) { return g_err.c_str(); }

extern "C" int cv_device_info(int device, char *name, int name_len, int32_t *cus) {
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (name && name_len > 0) {
        snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cus) *cus = prop.multiProcessorCount;
    return CV_OK;
}

// rcpn: every float with biased exponent 1..252 (2^-126 <= |b| < 2^126), both signs, against
// the correctly rounded 1/b; sqrtn: every float of [2^-24, 1] against sqrtf
#define CV_RCPN_PER_SIGN (252ull << 23)
#define CV_SQRTN_LO 0x33800000u           // 2^-24
#define CV_SQRTN_HI 0x3f800000u           // 1
__global__ void k_selftest_fastmath(unsigned long long *bad) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t nr = 2 * CV_RCPN_PER_SIGN, ns = (uint64_t)(CV_SQRTN_HI - CV_SQRTN_LO) + 1;
    unsigned long long cr = 0, cs = 0;
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nr + ns; u += stride) {
        if (u < nr) {
            const uint32_t sign = u >= CV_RCPN_PER_SIGN ? 0x80000000u : 0u;
            const float b = __uint_as_float(sign | (uint32_t)(u % CV_RCPN_PER_SIGN + (1ull << 23)));
            cr += __float_as_uint(rcpn(b)) != __float_as_uint(rcp(b));
        } else {
            const float y = __uint_as_float(CV_SQRTN_LO + (uint32_t)(u - nr));
            cs += __float_as_uint(sqrtn(y)) != __float_as_uint(sqrtf(y));
        }
    }
    if (cr) atomicAdd(&bad[0], cr);
    if (cs) atomicAdd(&bad[1], cs);
}

extern "C" int cv_selftest_fastmath(int device, uint64_t out[4]) {
    if (!out) return fail(CV_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(device));
    unsigned long long *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, 2 * sizeof(unsigned long long)));
    unsigned long long h[2] = {0, 0};
    hipError_t e = hipMemcpy(d, h, sizeof(h), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_selftest_fastmath, dim3(16384), dim3(256), 0, 0, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_selftest_fastmath: %s", hipGetErrorString(e));
    out[0] = h[0];
    out[1] = 2 * CV_RCPN_PER_SIGN;
    out[2] = h[1];
    out[3] = (uint64_t)(CV_SQRTN_HI - CV_SQRTN_LO) + 1;
    return CV_OK;
}

static StageP stage_params(const cv_handle *h, int step, float tau) {
    StageP s;
    s.step = step;
    s.tau = tau;
    s.r_tau = 1.0f / tau;
    s.r_taudz = 1.0f / (tau * h->prm.dz);
    s.old_time = tau - (float)(step - 1) * (float)h->prm.dt;
    s.r_step = 1.0f / (float)step;
    return s;
}

static float tau_at(const cv_handle *h, int64_t loop_plus_one) {
    return (float)(h->prm.tau0 + (double)loop_plus_one * h->prm.dt);
}

extern "C" int cv_create(const cv_params *p, const cv_eos *eos, int device, int viscous, cv_handle **out) {
    if (!p || !eos || !out) return fail(CV_ERR_ARG, "cv_create: null argument");
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(CV_ERR_ARG, "cv_create: grid must be >= 1 in each axis");
    if (!(p->dx > 0 && p->dy > 0 && p->dz > 0 && p->dt > 0 && p->tau0 > 0))
        return fail(CV_ERR_ARG, "cv_create: DX, DY, DZ, DT, TAU0 must be positive");
    if (!eos->rgba || eos->n_entries < 2) return fail(CV_ERR_ARG, "cv_create: EOS table missing");
    if (eos->kind == 0) {
        if ((int64_t)eos->nrow * eos->ncol != eos->n_entries || eos->ed_step <= 0)
            return fail(CV_ERR_ARG, "cv_create: row/col EOS needs nrow*ncol == n_entries and ed_step > 0");
    } else if (eos->kind == 1) {
        if (eos->log2j < 0 || eos->log2j > 16 || eos->kmax <= eos->kmin || eos->kmin < -126 || eos->kmax > 127 ||
            (int64_t)(eos->kmax - eos->kmin) * (1 << eos->log2j) + 1 != eos->n_entries)
            return fail(CV_ERR_ARG, "cv_create: octave EOS needs (kmax-kmin)*2^log2j+1 == n_entries");
    } else {
        return fail(CV_ERR_ARG, "cv_create: unknown EOS kind %d", eos->kind);
    }
    if ((p->flags & CV_FLAG_BULK) && !viscous) return fail(CV_ERR_ARG, "cv_create: bulk needs viscous");
    if ((int64_t)p->nx * p->ny * p->nz >= ((int64_t)1 << 30))
        return fail(CV_ERR_ARG, "cv_create: %lld cells exceed the 2^30 cells of one field plane",
                    (long long)p->nx * p->ny * p->nz);
    HIP_TRY(hipSetDevice(device));
    cv_handle *h = new cv_handle();
    h->prm = *p;
    h->device = device;
    h->viscous = viscous;
    h->bulk = (p->flags & CV_FLAG_BULK) ? 1 : 0;
    int64_t n = (int64_t)p->nx * p->ny * p->nz;
    h->n = n;
    PhysP &ph = h->phys;
    ph.nx = p->nx; ph.ny = p->ny; ph.nz = p->nz;
    ph.n = n; ph.sx = (int64_t)p->ny * p->nz; ph.sy = p->nz;
    ph.dx = p->dx; ph.dy = p->dy; ph.dz = p->dz; ph.dt = (float)p->dt;
    ph.r_dx = 1.0f / p->dx; ph.r_dy = 1.0f / p->dy; ph.r_dz = 1.0f / p->dz;
    ph.r_dt = 1.0f / ph.dt; ph.r_3 = 1.0f / 3.0f;
    ph.etaos_xmin = p->etaos_xmin; ph.etaos_ymin = p->etaos_ymin;
    ph.etaos_ls = p->etaos_left_slop; ph.etaos_rs = p->etaos_right_slop; ph.lam1 = p->lam1;
    ph.zmax = p->zetaos_max; ph.ztpeak = p->zetaos_tpeak; ph.zwidth = p->zetaos_width;
    ph.r_zwidth = 1.0f / p->zetaos_width;
    ph.fp64_vmin = p->fp64_root_vmin;
    ph.flags = p->flags;
    // 32-bit plane offsets of the tiled kernel: 10 planes of 4 n bytes below 2^32
    ph.nb = (40 * n < ((int64_t)1 << 32)) ? (uint32_t)(4 * n) : 0u;
    // the max(lambda, 1) shortcut needs every cs2 the lookup can return in [0, 0.5]: the
    // table's nodes (a lookup interpolates between two of them or scales none) and the
    // 1/3 of the fit above a row/col table
    float cs2max = -1e30f, cs2min = 1e30f;
    for (int64_t i = 0; i < eos->n_entries; i++) {
        const float c = eos->rgba[4 * i];
        cs2max = c > cs2max ? c : cs2max;
        cs2min = c < cs2min ? c : cs2min;
    }
    ph.lam_fast = (cs2max <= 0.5f && cs2min >= 0.0f) ? 1 : 0;
    EosP &E = h->eos;
    memset(&E, 0, sizeof(E));
    E.kind = eos->kind;
    if (eos->kind == 0) {
        E.nrow = eos->nrow; E.ncol = eos->ncol; E.num_ed = eos->num_ed;
        E.ed_start = eos->ed_start; E.ed_step = eos->ed_step;
        E.epr = eos->ed_step * eos->ncol;
        E.r_epr = 1.0f / E.epr;
        E.r_step = 1.0f / eos->ed_step;
        E.top = eos->ed_start + (float)eos->num_ed * eos->ed_step;
    } else {
        E.kmin = eos->kmin; E.kmax = eos->kmax; E.lj = eos->log2j;
        E.nlast = (eos->kmax - eos->kmin) << eos->log2j;
        E.lo_e = ldexpf(1.0f, eos->kmin); E.hi_e = ldexpf(1.0f, eos->kmax);
        E.r_lo = ldexpf(1.0f, -eos->kmin); E.r_hi = ldexpf(1.0f, -eos->kmax);
    }
    int rc = CV_OK;
    auto alloc = [&](float **ptr, int64_t nfl) -> bool {
        if (hipMalloc((void **)ptr, sizeof(float) * nfl) != hipSuccess) return false;
        return hipMemset(*ptr, 0, sizeof(float) * nfl) == hipSuccess;
    };
    bool ok = hipMalloc((void **)&h->d_eos, sizeof(float4) * eos->n_entries) == hipSuccess &&
              hipMemcpy(h->d_eos, eos->rgba, sizeof(float4) * eos->n_entries, hipMemcpyHostToDevice) == hipSuccess;
    E.tab = h->d_eos;
    for (int k = 0; ok && k < 3; k++) ok = alloc(&h->ev[k], 4 * n);
    ok = ok && alloc(&h->Src, 4 * n);
    if (viscous) {
        for (int k = 0; ok && k < 3; k++) ok = alloc(&h->pi[k], 10 * n);
        ok = ok && alloc(&h->IS_src, 10 * n) && alloc(&h->udiff, 4 * n) && alloc(&h->good, n);
        if (ok) {
            std::vector<float> ones(n, 1.0f);
            ok = hipMemcpy(h->good, ones.data(), sizeof(float) * n, hipMemcpyHostToDevice) == hipSuccess;
        }
        if (h->bulk) {
            for (int k = 0; ok && k < 3; k++) ok = alloc(&h->bk[k], n);
            ok = ok && alloc(&h->bsrc, n);
        }
        // deferred FP64 root solves of the tiled stage kernels: n entries always suffice
        if (ok && (p->flags & CV_FLAG_FP64_ROOT))
            ok = hipMalloc((void **)&h->d_fix, sizeof(FixEntry) * n) == hipSuccess &&
                 hipMalloc((void **)&h->d_fix_ctr, sizeof(unsigned)) == hipSuccess &&
                 hipMemset(h->d_fix_ctr, 0, sizeof(unsigned)) == hipSuccess;
    }
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) h->cus = prop.multiProcessorCount;
    }
    if (!ph.nb) h->kernel = 0;     // the tiled kernel needs 32-bit plane offsets
    else if (!march_ok(h)) h->kernel = 1;
    if (ok && (rc = default_tile(h)) != CV_OK) {
        cv_destroy(h);
        return rc;
    }
    ok = ok && hipMalloc((void **)&h->d_max, sizeof(unsigned int)) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreate(&h->ev_a) == hipSuccess && hipEventCreate(&h->ev_b) == hipSuccess;
    if (!ok) {
        rc = fail(CV_ERR_HIP, "cv_create: device allocation failed (%lld cells): %s", (long long)n,
                  hipGetErrorString(hipGetLastError()));
        cv_destroy(h);
        return rc;
    }
    *out = h;
    return CV_OK;
}

extern "C" int cv_destroy(cv_handle *h) {
    if (!h) return CV_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (int k = 0; k < 3; k++) { (void)hipFree(h->ev[k]); (void)hipFree(h->pi[k]); (void)hipFree(h->bk[k]); }
    (void)hipFree(h->Src); (void)hipFree(h->IS_src); (void)hipFree(h->bsrc);
    (void)hipFree(h->udiff); (void)hipFree(h->good);
    (void)hipFree(h->d_eos); (void)hipFree(h->d_max); (void)hipFree(h->d_esum);
    (void)hipFree(h->d_fix); (void)hipFree(h->d_fix_ctr);
    (void)hipFree(h->sf_ev); (void)hipFree(h->sf_pi); (void)hipFree(h->sf_bk);
    (void)hipFree(h->sf_out); (void)hipFree(h->sf_txyz); (void)hipFree(h->sf_pisf); (void)hipFree(h->sf_bksf);
    (void)hipFree(h->sf_om); (void)hipFree(h->om[0]); (void)hipFree(h->om[1]);
    (void)hipFree(h->sf_count); (void)hipFree(h->sf_list); (void)hipFree(h->sf_nlist); (void)hipFree(h->d_slice);
    (void)hipFree(h->seed_buf); (void)hipFree(h->iters);
    if (h->ev_a) (void)hipEventDestroy(h->ev_a);
    if (h->ev_b) (void)hipEventDestroy(h->ev_b);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return CV_OK;
}

// field id -> (device pointer, number of components)
static int field_ptr(cv_handle *h, int f, float **ptr, int *ncomp) {
    switch (f) {
    case CV_EV0: case CV_EV1: case CV_EV2: *ptr = h->ev[f - CV_EV0]; *ncomp = 4; break;
    case CV_PI0: case CV_PI1: case CV_PI2: *ptr = h->pi[f - CV_PI0]; *ncomp = 10; break;
    case CV_BULK0: case CV_BULK1: case CV_BULK2: *ptr = h->bk[f - CV_BULK0]; *ncomp = 1; break;
    case CV_SRC: *ptr = h->Src; *ncomp = 4; break;
    case CV_IS_SRC: *ptr = h->IS_src; *ncomp = 10; break;
    case CV_UDIFF: *ptr = h->udiff; *ncomp = 4; break;
    case CV_GOODCELL: *ptr = h->good; *ncomp = 1; break;
    case CV_ROOT_CONS: *ptr = h->seed_buf; *ncomp = 4; break;
    case CV_ROOT_SEED: *ptr = h->seed_buf ? h->seed_buf + 9 * h->n : nullptr; *ncomp = 1; break;
    case CV_OMEGA0: case CV_OMEGA1: *ptr = h->om[f - CV_OMEGA0]; *ncomp = CV_VORT_N; break;
    default: return fail(CV_ERR_ARG, "unknown field id %d", f);
    }
    if (!*ptr) return fail(CV_ERR_STATE, "field %d is not allocated for this handle (ideal or no bulk)", f);
    return CV_OK;
}

extern "C" int cv_set_field(cv_handle *h, int f, const float *host, int64_t nfloats) {
    if (!h || !host) return fail(CV_ERR_ARG, "cv_set_field: null argument");
    float *d; int nc;
    int rc = field_ptr(h, f, &d, &nc);
    if (rc) return rc;
    if (nfloats != nc * h->n) return fail(CV_ERR_ARG, "cv_set_field: expected %lld floats, got %lld",
                                          (long long)(nc * h->n), (long long)nfloats);
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    std::vector<float> soa((size_t)nfloats);
    for (int64_t I = 0; I < h->n; I++)
        for (int c = 0; c < nc; c++) soa[(size_t)(c * h->n + I)] = host[I * nc + c];
    HIP_TRY(hipMemcpy(d, soa.data(), sizeof(float) * nfloats, hipMemcpyHostToDevice));
    return CV_OK;
}

extern "C" int cv_get_field(cv_handle *h, int f, float *host, int64_t nfloats) {
    if (!h || !host) return fail(CV_ERR_ARG, "cv_get_field: null argument");
    float *d; int nc;
    int rc = field_ptr(h, f, &d, &nc);
    if (rc) return rc;
    if (nfloats != nc * h->n) return fail(CV_ERR_ARG, "cv_get_field: expected %lld floats, got %lld",
                                          (long long)(nc * h->n), (long long)nfloats);
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    std::vector<float> soa((size_t)nfloats);
    HIP_TRY(hipMemcpy(soa.data(), d, sizeof(float) * nfloats, hipMemcpyDeviceToHost));
    for (int64_t I = 0; I < h->n; I++)
        for (int c = 0; c < nc; c++) host[I * nc + c] = soa[(size_t)(c * h->n + I)];
    return CV_OK;
}

extern "C" int cv_copy_field(cv_handle *h, int dst, int src) {
    if (!h) return fail(CV_ERR_ARG, "cv_copy_field: null handle");
    float *a, *b; int na, nb_;
    int rc = field_ptr(h, dst, &a, &na);
    if (rc) return rc;
    rc = field_ptr(h, src, &b, &nb_);
    if (rc) return rc;
    if (na != nb_) return fail(CV_ERR_ARG, "cv_copy_field: component counts differ");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipMemcpyAsync(a, b, sizeof(float) * na * h->n, hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CV_OK;
}

static dim3 grid1d(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

extern "C" int cv_gather(cv_handle *h, int f, const int64_t *cells, int64_t ncells, float *host) {
    if (!h || (ncells > 0 && (!cells || !host))) return fail(CV_ERR_ARG, "cv_gather: null argument");
    if (ncells < 0 || ncells > ((int64_t)1 << 31)) return fail(CV_ERR_ARG, "cv_gather: bad cell count");
    float *d; int nc;
    int rc = field_ptr(h, f, &d, &nc);
    if (rc) return rc;
    for (int64_t q = 0; q < ncells; q++)
        if (cells[q] < 0 || cells[q] >= h->n)
            return fail(CV_ERR_ARG, "cv_gather: cell %lld out of [0, %lld)", (long long)cells[q], (long long)h->n);
    if (ncells == 0) return CV_OK;
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    int64_t *d_idx = nullptr;
    float *d_out = nullptr;
    hipError_t e = hipMalloc((void **)&d_idx, sizeof(int64_t) * ncells);
    if (e == hipSuccess) e = hipMalloc((void **)&d_out, sizeof(float) * nc * ncells);
    if (e == hipSuccess) e = hipMemcpy(d_idx, cells, sizeof(int64_t) * ncells, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_gather, grid1d(ncells), dim3(256), 0, h->stream, h->n, nc, d, d_idx, ncells, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(host, d_out, sizeof(float) * nc * ncells, hipMemcpyDeviceToHost);
    (void)hipFree(d_out);
    (void)hipFree(d_idx);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_gather: %s", hipGetErrorString(e));
    return CV_OK;
}

extern "C" int cv_visc_initialize(cv_handle *h, float tau) {
    if (!h || !h->viscous) return fail(CV_ERR_STATE, "cv_visc_initialize: viscous handle required");
    HIP_TRY(hipSetDevice(h->device));
    hipLaunchKernelGGL(k_visc_init, grid1d(h->n), dim3(256), 0, h->stream, h->phys, h->eos, tau, 1.0f / tau,
                       h->ev[1], h->pi[1], h->good, h->udiff);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CV_OK;
}

// ---- launch helpers (queue only) --------------------------------------------
static int launch_ideal_seeded(cv_handle *h, const StageP &st, float *ev_st, float *ev1, float *ev_out);

static int launch_ideal(cv_handle *h, int step, float tau, float *ev_st, float *ev1, float *ev_out) {
    StageP st = stage_params(h, step, tau);
    if (h->seed_mlp) return launch_ideal_seeded(h, st, ev_st, ev1, ev_out);
    // x-marching tiles with the marching stage kernel's tile and chunks; the test flags, the
    // FP64 root and the iteration counts take the per-cell kernel
    if (h->kernel == 2 && h->phys.nb && !(h->prm.flags & (CV_FLAG_RIEMANN_TEST | CV_FLAG_FP64_ROOT)) &&
        !h->track_iters && !h->jet.on) {
        dim3 g((unsigned)h->tile.nwork), b(CV_ITILE_THREADS);
        if (h->eos.kind == 1)
            hipLaunchKernelGGL(k_ideal_tiled<1>, g, b, 0, h->stream, h->phys, h->eos, st, h->tile, ev_st, ev1,
                               h->Src, ev_out);
        else
            hipLaunchKernelGGL(k_ideal_tiled<0>, g, b, 0, h->stream, h->phys, h->eos, st, h->tile, ev_st, ev1,
                               h->Src, ev_out);
        HIP_TRY(hipGetLastError());
        return CV_OK;
    }
    hipLaunchKernelGGL(k_ideal_stage, grid1d(h->n), dim3(256), 0, h->stream, h->phys, h->eos, st,
                       ev_st, ev1, h->Src, ev_out, nullptr, nullptr, h->track_iters ? h->iters : nullptr, h->jet);
    HIP_TRY(hipGetLastError());
    return CV_OK;
}

static int launch_is(cv_handle *h, int step, float tau, float *ev_st, float *pi_st, float *bk_st,
                     float *ev1, float *ev2, float *pi1, float *bk1, float *pi_out, float *bk_out) {
    StageP st = stage_params(h, step, tau);
    if (h->bulk)
        hipLaunchKernelGGL(k_is_stage<true>, grid1d(h->n), dim3(256), 0, h->stream, h->phys, h->eos, st,
                           ev_st, pi_st, bk_st, ev1, ev2, pi1, bk1, h->udiff, h->IS_src, h->bsrc,
                           pi_out, bk_out, h->good);
    else
        hipLaunchKernelGGL(k_is_stage<false>, grid1d(h->n), dim3(256), 0, h->stream, h->phys, h->eos, st,
                           ev_st, pi_st, bk_st, ev1, ev2, pi1, bk1, h->udiff, h->IS_src, h->bsrc,
                           pi_out, bk_out, h->good);
    HIP_TRY(hipGetLastError());
    return CV_OK;
}

static int launch_kt(cv_handle *h, int step, float tau, float *ev_st, float *pi_st, float *bk_st,
                     float *ev1, float *pim0, float *bk0, float *pinew, float *bknew, float *ev_out,
                     float *udiff_out) {
    StageP st = stage_params(h, step, tau);
    if (h->bulk)
        hipLaunchKernelGGL(k_kt_stage<true>, grid1d(h->n), dim3(256), 0, h->stream, h->phys, h->eos, st,
                           ev_st, pi_st, bk_st, ev1, pim0, bk0, pinew, bknew, h->Src, ev_out, udiff_out, h->jet);
    else
        hipLaunchKernelGGL(k_kt_stage<false>, grid1d(h->n), dim3(256), 0, h->stream, h->phys, h->eos, st,
                           ev_st, pi_st, bk_st, ev1, pim0, bk0, pinew, bknew, h->Src, ev_out, udiff_out, h->jet);
    HIP_TRY(hipGetLastError());
    return CV_OK;
}

// the seed pass's packed BF16 operands from the network's current parameters (on the stage's
// stream, so that training between stages is always seen)
static void mlp_pack(const cv_mlp *m, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_pack_t, dim3((MLP_PK_THREADS + 255) / 256), dim3(256), 0, s, m->P, m->pk);
}

// second half of a split viscous stage: the learned seed (when set) from the conserved
// input the stage wrote, then the root solve and velocities
static int launch_visc_root(cv_handle *h, int step, const float *ev1, const float *bk_out, float *ev_out,
                            float *udiff_out) {
    const int64_t n = h->n;
    const float *cons = h->seed_buf;
    float *seed = nullptr;
    if (h->seed_mlp) {
        seed = h->seed_buf + 9 * n;
        const int64_t per_blk = 16 * MLP_WAVES * MLP_SEED_TILES;
        const int64_t nblk = (n + per_blk - 1) / per_blk;
        mlp_pack(h->seed_mlp, h->stream);
        hipLaunchKernelGGL(k_mlp_seed_cons, dim3((unsigned)nblk), dim3(64 * MLP_WAVES), 0, h->stream, h->seed_mlp->P,
                           h->seed_mlp->pk, cons, n, seed);
    }
    int *it = h->track_iters ? h->iters : nullptr;
#define CV_ROOT_LAUNCH(B, S, K) \
    hipLaunchKernelGGL((k_visc_root<B, S, K>), grid1d(n), dim3(256), 0, h->stream, h->phys, h->eos, cons, bk_out, \
                       ev1, seed, ev_out, udiff_out, it)
#define CV_ROOT_SEL(B, S) do { if (h->eos.kind == 1) CV_ROOT_LAUNCH(B, S, 1); else CV_ROOT_LAUNCH(B, S, 0); } while (0)
    if (h->bulk) {
        if (step == 1) CV_ROOT_SEL(true, 1); else CV_ROOT_SEL(true, 2);
    } else {
        if (step == 1) CV_ROOT_SEL(false, 1); else CV_ROOT_SEL(false, 2);
    }
#undef CV_ROOT_SEL
#undef CV_ROOT_LAUNCH
    HIP_TRY(hipGetLastError());
    return CV_OK;
}

// fused IS + KT stage of the step path (k_visc_stage)
static int launch_stage(cv_handle *h, int step, float tau, float *ev_st, float *pi_st, float *bk_st,
                        float *ev1, float *ev2, float *pi1, float *bk1, float *pi_out, float *bk_out,
                        float *ev_out, float *udiff_out) {
    StageP st = stage_params(h, step, tau);
    // the jet source runs in the per-cell kernel
    if ((h->kernel == 1 || h->kernel == 2) && !h->jet.on) {
        const bool march = h->kernel == 2;
        dim3 g(march ? (unsigned)h->tile.nwork : (unsigned)h->tile_grid), b(CV_TILE_THREADS);
        const bool tests = (h->prm.flags & (CV_FLAG_RIEMANN_TEST | CV_FLAG_GUBSER_VISC_TEST | CV_FLAG_PIMN_OMEGA |
                                            CV_FLAG_OMEGA_OMEGA)) != 0;
        const bool fp64 = (h->prm.flags & CV_FLAG_FP64_ROOT) != 0;
        const int ek = h->eos.kind;
        // the learned seed runs inside the stage (the whole wave evaluates the network before
        // the root solve) on the x-marching kernel with an octave table, unless the split form
        // is asked for; otherwise a seeded stage runs split (the same fields either way)
        const bool fused_seed = h->seed_mlp && !h->root_split && march && ek == 1 && !tests;
        const bool split = h->root_split || (h->seed_mlp && !fused_seed);
        StagePtrs sp = {ev_st, pi_st, bk_st, ev1, ev2, pi1, bk1, h->udiff, h->IS_src,
                        h->bsrc, h->Src, pi_out, bk_out, ev_out, h->good, udiff_out,
                        split ? h->seed_buf : nullptr, h->d_fix, h->d_fix_ctr,
                        fused_seed ? h->seed_mlp->P : nullptr, fused_seed ? h->seed_mlp->pk : nullptr,
                        fused_seed && h->track_iters ? h->seed_buf + 9 * h->n : nullptr,
                        fused_seed && h->track_iters ? h->iters : nullptr};
#define CV_TILED_ARGS h->phys, h->eos, st, h->tile, sp
        // optional branches: none, the FP64 root alone, the split alone (k_visc_root solves the
        // FP64 cells of a split stage), or all of them (test flags)
        const int spec = tests ? CV_SPEC_ALL : (split ? CV_SPEC_SPLIT : (fp64 ? CV_SPEC_FP64 : 0));
        // FP64 cells are deferred to k_fp64_fix unless the split stage solves every cell later
        const bool defer = fp64 && !split;
        if (defer) hipLaunchKernelGGL(k_fix_reset, dim3(1), dim3(64), 0, h->stream, h->d_fix_ctr);
        if (fused_seed) mlp_pack(h->seed_mlp, h->stream);
#define CV_TILED_LAUNCH2(B, S, K, X, M) hipLaunchKernelGGL((k_visc_tiled<B, S, K, X, M>), g, b, 0, h->stream, CV_TILED_ARGS)
#define CV_TILED_LAUNCH(B, S, K, X) \
        do { if (march) CV_TILED_LAUNCH2(B, S, K, X, true); else CV_TILED_LAUNCH2(B, S, K, X, false); } while (0)
#define CV_TILED_SEL_K(B, S, K)                                                                  \
        do { if (spec == CV_SPEC_ALL) CV_TILED_LAUNCH(B, S, K, CV_SPEC_ALL);                         \
             else if (spec == CV_SPEC_FP64) CV_TILED_LAUNCH(B, S, K, CV_SPEC_FP64);                  \
             else if (spec == CV_SPEC_SPLIT) CV_TILED_LAUNCH(B, S, K, CV_SPEC_SPLIT);                \
             else CV_TILED_LAUNCH(B, S, K, 0); } while (0)
#define CV_TILED_SEL(B, S) \
        if (ek == 1) CV_TILED_SEL_K(B, S, 1); else CV_TILED_SEL_K(B, S, 0);
#define CV_SEED_LAUNCH(B, S)                                                                                     \
        do { if (fp64) CV_TILED_LAUNCH2(B, S, 1, CV_SPEC_SEED | CV_SPEC_FP64, true);                              \
             else CV_TILED_LAUNCH2(B, S, 1, CV_SPEC_SEED, true); } while (0)
        if (fused_seed) {
            if (h->bulk) { if (step == 1) CV_SEED_LAUNCH(true, 1); else CV_SEED_LAUNCH(true, 2); }
            else { if (step == 1) CV_SEED_LAUNCH(false, 1); else CV_SEED_LAUNCH(false, 2); }
        } else if (h->bulk) {
            if (step == 1) { CV_TILED_SEL(true, 1) } else { CV_TILED_SEL(true, 2) }
        } else {
            if (step == 1) { CV_TILED_SEL(false, 1) } else { CV_TILED_SEL(false, 2) }
        }
#undef CV_SEED_LAUNCH
#undef CV_TILED_SEL
#undef CV_TILED_SEL_K
#undef CV_TILED_LAUNCH
#undef CV_TILED_LAUNCH2
#undef CV_TILED_ARGS
        HIP_TRY(hipGetLastError());
        if (split) return launch_visc_root(h, step, ev1, bk_out, ev_out, udiff_out);
        if (defer) {
            dim3 fg(4 * (h->cus > 0 ? h->cus : 256)), fb(256);
#define CV_FIX_LAUNCH(B, S, K) hipLaunchKernelGGL((k_fp64_fix<B, S, K>), fg, fb, 0, h->stream, h->phys, h->eos, \
                                                  h->d_fix, h->d_fix_ctr, ev1, ev_out, udiff_out)
#define CV_FIX_SEL(B, S) do { if (ek == 1) CV_FIX_LAUNCH(B, S, 1); else CV_FIX_LAUNCH(B, S, 0); } while (0)
            if (h->bulk) { if (step == 1) CV_FIX_SEL(true, 1); else CV_FIX_SEL(true, 2); }
            else { if (step == 1) CV_FIX_SEL(false, 1); else CV_FIX_SEL(false, 2); }
#undef CV_FIX_SEL
#undef CV_FIX_LAUNCH
            HIP_TRY(hipGetLastError());
        }
        return CV_OK;
    }
    if ((h->root_split || h->seed_mlp) && h->jet.on)
        return fail(CV_ERR_UNSUPPORTED, "the jet source does not run with the split or seeded root solve");
    if (h->root_split || h->seed_mlp)
        return fail(CV_ERR_UNSUPPORTED, "the split or seeded root solve needs a tiled stage kernel (kind 1 or 2)");
    dim3 g = grid1d(h->n), b(256);
#define CV_STAGE_ARGS h->phys, h->eos, st, ev_st, pi_st, bk_st, ev1, ev2, pi1, bk1, h->udiff, h->IS_src, \
                      h->bsrc, h->Src, pi_out, bk_out, ev_out, h->good, udiff_out, h->jet
    if (h->bulk) {
        if (step == 1) hipLaunchKernelGGL((k_visc_stage<true, 1>), g, b, 0, h->stream, CV_STAGE_ARGS);
        else hipLaunchKernelGGL((k_visc_stage<true, 2>), g, b, 0, h->stream, CV_STAGE_ARGS);
    } else {
        if (step == 1) hipLaunchKernelGGL((k_visc_stage<false, 1>), g, b, 0, h->stream, CV_STAGE_ARGS);
        else hipLaunchKernelGGL((k_visc_stage<false, 2>), g, b, 0, h->stream, CV_STAGE_ARGS);
    }
#undef CV_STAGE_ARGS
    HIP_TRY(hipGetLastError());
    return CV_OK;
}

// one RK2 step of the evolve loop; stage 2 writes ev_{n+1}, pi_{n+1} into the slot-0
// buffers, then slots 0 and 1 swap (see cv_visc_steps_async)
static int launch_visc_step(cv_handle *h, int64_t n, hipEvent_t mid = nullptr) {
    float t1 = tau_at(h, n), t2 = tau_at(h, n + 1);
    float **E = h->ev, **P = h->pi, **B = h->bk;
    int rc = launch_stage(h, 1, t1, E[1], P[1], B[1], E[1], E[2], P[1], B[1], P[2], B[2], E[2], nullptr);
    if (rc) return rc;
    if (mid) HIP_TRY(hipEventRecord(mid, h->stream));
    rc = launch_stage(h, 2, t2, E[2], P[2], B[2], E[1], E[2], P[1], B[1], P[0], B[0], E[0], h->udiff);
    if (rc) return rc;
    float *t;
    t = E[0]; E[0] = E[1]; E[1] = t;
    t = P[0]; P[0] = P[1]; P[1] = t;
    t = B[0]; B[0] = B[1]; B[1] = t;
    return CV_OK;
}

static int check_step(int step) {
    if (step != 1 && step != 2) return fail(CV_ERR_ARG, "step must be 1 or 2, got %d", step);
    return CV_OK;
}

extern "C" int cv_ideal_stage(cv_handle *h, int step, float tau) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    int rc = check_step(step);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(h->device));
    rc = launch_ideal(h, step, tau, h->ev[step], h->ev[1], h->ev[3 - step]);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CV_OK;
}

extern "C" int cv_is_stage(cv_handle *h, int step, float tau) {
    if (!h || !h->viscous) return fail(CV_ERR_STATE, "cv_is_stage: viscous handle required");
    int rc = check_step(step);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(h->device));
    rc = launch_is(h, step, tau, h->ev[step], h->pi[step], h->bk[step], h->ev[1], h->ev[2], h->pi[1], h->bk[1],
                   h->pi[3 - step], h->bk[3 - step]);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CV_OK;
}

extern "C" int cv_visc_stage(cv_handle *h, int step, float tau) {
    if (!h || !h->viscous) return fail(CV_ERR_STATE, "cv_visc_stage: viscous handle required");
    int rc = check_step(step);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(h->device));
    rc = launch_kt(h, step, tau, h->ev[step], h->pi[step], h->bk[step], h->ev[1], h->pi[0], h->bk[0],
                   h->pi[3 - step], h->bk[3 - step], h->ev[3 - step], nullptr);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CV_OK;
}

extern "C" int cv_update_udiff(cv_handle *h, int a, int b) {
    if (!h || !h->viscous) return fail(CV_ERR_STATE, "cv_update_udiff: viscous handle required");
    if (a < CV_EV0 || a > CV_EV2 || b < CV_EV0 || b > CV_EV2) return fail(CV_ERR_ARG, "ev field ids expected");
    HIP_TRY(hipSetDevice(h->device));
    hipLaunchKernelGGL(k_udiff, grid1d(h->n), dim3(256), 0, h->stream, h->phys, h->ev[a], h->ev[b], h->udiff);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CV_OK;
}

extern "C" int cv_ideal_steps_async(cv_handle *h, int64_t loop0, int64_t nsteps) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    if (nsteps < 0 || loop0 < 0) return fail(CV_ERR_ARG, "negative loop index or step count");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipEventRecord(h->ev_a, h->stream));
    for (int64_t n = loop0; n < loop0 + nsteps; n++) {
        int rc = launch_ideal(h, 1, tau_at(h, n), h->ev[1], h->ev[1], h->ev[2]);
        if (rc) return rc;
        rc = launch_ideal(h, 2, tau_at(h, n + 1), h->ev[2], h->ev[1], h->ev[1]);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(h->ev_b, h->stream));
    h->timed = true;
    return CV_OK;
}

// One step of CLVisc.evolve's body with buffer rotation instead of the two D2D copies:
// stage 2 writes ev_{n+1}, pi_{n+1} into the slot-0 buffers, then slots 0 and 1 swap, which
// leaves d_ev[0] = ev_n and d_pi[0] = pi_n exactly as the copies of visc.py:527-532 do.
extern "C" int cv_visc_steps_async(cv_handle *h, int64_t loop0, int64_t nsteps) {
    if (!h || !h->viscous) return fail(CV_ERR_STATE, "cv_visc_steps_async: viscous handle required");
    if (nsteps < 0 || loop0 < 0) return fail(CV_ERR_ARG, "negative loop index or step count");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipEventRecord(h->ev_a, h->stream));
    for (int64_t n = loop0; n < loop0 + nsteps; n++) {
        int rc = launch_visc_step(h, n);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(h->ev_b, h->stream));
    h->timed = true;
    return CV_OK;
}

extern "C" int cv_profile_visc_step(cv_handle *h, int64_t loop0, float *ms, int nms, int32_t *nlaunch) {
    if (!h || !h->viscous) return fail(CV_ERR_STATE, "cv_profile_visc_step: viscous handle required");
    if (!ms || !nlaunch || loop0 < 0) return fail(CV_ERR_ARG, "cv_profile_visc_step: bad argument");
    HIP_TRY(hipSetDevice(h->device));
    hipEvent_t e0, e1, e2;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventCreate(&e2));
    HIP_TRY(hipEventRecord(e0, h->stream));
    int rc = launch_visc_step(h, loop0, e1);
    if (!rc) {
        HIP_TRY(hipEventRecord(e2, h->stream));
        HIP_TRY(hipEventSynchronize(e2));
        float a = 0.0f, b = 0.0f;
        HIP_TRY(hipEventElapsedTime(&a, e0, e1));
        HIP_TRY(hipEventElapsedTime(&b, e1, e2));
        if (nms > 0) ms[0] = a;
        if (nms > 1) ms[1] = b;
        *nlaunch = 2;
    }
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(e2);
    return rc;
}

extern "C" int cv_set_jet_source(cv_handle *h, int on, float direction, int64_t start_pos_index) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    if (!on) { h->jet = JetP{}; return CV_OK; }
    if (on < 0 || on > 3) return fail(CV_ERR_ARG, "on must be a mask of the RK stages (1, 2 or 3)");
    if (start_pos_index < 0 || start_pos_index >= h->n) return fail(CV_ERR_ARG, "start_pos_index outside the grid");
    if (h->root_split || (h->viscous && h->seed_mlp))
        return fail(CV_ERR_UNSUPPORTED, "the jet source does not run with the split or seeded viscous root solve");
    const int64_t nyz = (int64_t)h->prm.ny * h->prm.nz;
    const int i0 = (int)(start_pos_index / nyz);
    const int j0 = (int)((start_pos_index - i0 * nyz) / h->prm.nz);
    const int k0 = (int)(start_pos_index - i0 * nyz - (int64_t)j0 * h->prm.nz);
    JetP j;
    j.on = on;
    j.c = (float)cos((double)direction);
    j.sn = (float)sin((double)direction);
    j.x0 = (float)(i0 - h->prm.nx / 2) * h->phys.dx;
    j.y0 = (float)(j0 - h->prm.ny / 2) * h->phys.dy;
    j.z0 = (float)(k0 - h->prm.nz / 2) * h->phys.dz;
    j.c3 = (float)pow(sqrt(2.0 * M_PI), 3.0);
    j.tau0 = (float)h->prm.tau0;
    h->jet = j;
    return CV_OK;
}

extern "C" int cv_set_stage_kernel(cv_handle *h, int kind, int ty, int tz) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    if (kind < 0 || kind > 2) return fail(CV_ERR_ARG, "kernel kind must be 0 (per-cell), 1 (tiled) or 2 (tiled x-march)");
    if (kind >= 1 && !h->phys.nb)
        return fail(CV_ERR_UNSUPPORTED, "tiled kernel needs 40 * cells < 2^32 (%lld cells)", (long long)h->n);
    if (kind == 2 && !march_ok(h))
        return fail(CV_ERR_UNSUPPORTED, "x-march kernel: octave EOS tables of at most %d nodes (this one: %d)",
                    CV_PTAB_LDS, h->eos.nlast + 1);
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (kind >= 1) {
        int rc = (ty > 0 && tz > 0) ? set_tile(h, ty, tz) : default_tile(h);
        if (rc) return rc;
    }
    h->kernel = kind;
    return CV_OK;
}

extern "C" int cv_set_march_chunk(cv_handle *h, int xc) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    if (xc < 0) return fail(CV_ERR_ARG, "march chunk must be >= 0 (0: automatic)");
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->march_xc = xc;
    return set_tile(h, h->tile.ty, h->tile.tz);
}

extern "C" int cv_set_xcd_map(cv_handle *h, int nxcd) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    if (nxcd < 0 || nxcd > 64) return fail(CV_ERR_ARG, "cv_set_xcd_map: nxcd in [0, 64]");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->xcd = nxcd;
    return set_tile(h, h->tile.ty, h->tile.tz);
}

extern "C" int cv_get_march_chunk(cv_handle *h, int32_t *xc, int32_t *nchunks) {
    if (!h || !xc || !nchunks) return fail(CV_ERR_ARG, "null argument");
    *xc = h->tile.xc; *nchunks = h->tile.nchunks;
    return CV_OK;
}

extern "C" int cv_get_stage_kernel(cv_handle *h, int32_t *kind, int32_t *ty, int32_t *tz) {
    if (!h || !kind || !ty || !tz) return fail(CV_ERR_ARG, "null argument");
    *kind = h->kernel; *ty = h->tile.ty; *tz = h->tile.tz;
    return CV_OK;
}

extern "C" int cv_max_ed(cv_handle *h, int which, float *out) {
    if (!h || !out) return fail(CV_ERR_ARG, "null argument");
    if (which < 0 || which > 2) return fail(CV_ERR_ARG, "which must be 0, 1 or 2");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipMemsetAsync(h->d_max, 0, sizeof(unsigned int), h->stream));
    int64_t blocks = (h->n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_max_ed, dim3((unsigned)blocks), dim3(256), 0, h->stream, h->n, h->ev[which], h->d_max);
    HIP_TRY(hipGetLastError());
    unsigned int u = 0;
    HIP_TRY(hipMemcpyAsync(&u, h->d_max, sizeof(u), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    memcpy(out, &u, sizeof(float));
    return CV_OK;
}

extern "C" int cv_pizz_o_ep(cv_handle *h, int which, float *out) {
    if (!h || !out) return fail(CV_ERR_ARG, "null argument");
    if (which < 0 || which > 2) return fail(CV_ERR_ARG, "which must be 0, 1 or 2");
    if (!h->viscous) return fail(CV_ERR_UNSUPPORTED, "cv_pizz_o_ep: the handle has no pi^{mu nu}");
    HIP_TRY(hipSetDevice(h->device));
    float *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(float) * h->prm.nz));
    hipLaunchKernelGGL(k_pizz_o_ep, dim3((unsigned)h->prm.nz), dim3(64), 0, h->stream, h->phys, h->eos, h->ev[which],
                       h->pi[which], d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, sizeof(float) * h->prm.nz, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hipError_t e2 = hipFree(d);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_pizz_o_ep: %s", hipGetErrorString(e));
    if (e2 != hipSuccess) return fail(CV_ERR_HIP, "cv_pizz_o_ep: %s", hipGetErrorString(e2));
    return CV_OK;
}

extern "C" int cv_energy_sum(cv_handle *h, int which, double tau, double *out) {
    if (!h || !out) return fail(CV_ERR_ARG, "null argument");
    if (which < 0 || which > 2) return fail(CV_ERR_ARG, "which must be 0, 1 or 2");
    if (!(tau > 0.0)) return fail(CV_ERR_ARG, "cv_energy_sum: tau must be positive");
    HIP_TRY(hipSetDevice(h->device));
    if (!h->d_esum) HIP_TRY(hipMalloc((void **)&h->d_esum, sizeof(double) * (2 * CV_ESUM_BLOCKS + 2)));
    const float *pi = h->viscous ? h->pi[which] : nullptr;
    const float *bk = h->bulk ? h->bk[which] : nullptr;
    hipLaunchKernelGGL(k_energy_sum, dim3(CV_ESUM_BLOCKS), dim3(256), 0, h->stream, h->phys, h->eos, h->ev[which], pi,
                       bk, tau, h->d_esum);
    const double dV = (double)h->prm.dx * (double)h->prm.dy * (double)h->prm.dz;
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, h->stream, h->d_esum, CV_ESUM_BLOCKS, dV,
                       h->d_esum + 2 * CV_ESUM_BLOCKS);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, h->d_esum + 2 * CV_ESUM_BLOCKS, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CV_OK;
}

// Blocks of the slice-sum pass: one per group of (i, j) rows, at most 1024.
static int slice_blocks(const cv_handle *h) {
    return (int)std::min<int64_t>(1024, (int64_t)h->prm.nx * h->prm.ny);
}

// k_slice_sums into part (nb x nz x CV_SLICE_Q doubles), k_slice_final into sums (nz x
// CV_SLICE_Q doubles), on the handle's stream
static void launch_slice_sums(cv_handle *h, int which, double *part, double *sums) {
    const int nz = h->prm.nz;
    const int64_t nrow = (int64_t)h->prm.nx * h->prm.ny;
    const int nb = slice_blocks(h);
    const int64_t rpb = (nrow + nb - 1) / nb;
    const size_t nq = (size_t)nz * CV_SLICE_Q;
    hipLaunchKernelGGL(k_slice_sums, dim3(nb), dim3(128), 0, h->stream, h->phys, h->eos, h->ev[which], rpb, part);
    hipLaunchKernelGGL(k_slice_final, dim3((unsigned)nq), dim3(256), 0, h->stream, part, nb, nz, sums);
}

extern "C" int cv_slice_sums(cv_handle *h, int which, double *out) {
    if (!h || !out) return fail(CV_ERR_ARG, "null argument");
    if (which < 0 || which > 2) return fail(CV_ERR_ARG, "which must be 0, 1 or 2");
    HIP_TRY(hipSetDevice(h->device));
    const int nb = slice_blocks(h);
    const size_t nq = (size_t)h->prm.nz * CV_SLICE_Q;
    // nb and nz are fixed by the grid, so one buffer per handle serves every call
    if (!h->d_slice && hipMalloc((void **)&h->d_slice, sizeof(double) * (nq * (size_t)nb + nq)) != hipSuccess) {
        h->d_slice = nullptr;
        return fail(CV_ERR_HIP, "cv_slice_sums: allocation failed");
    }
    double *d = h->d_slice;
    launch_slice_sums(h, which, d, d + nq * (size_t)nb);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d + nq * (size_t)nb, sizeof(double) * nq, hipMemcpyDeviceToHost,
                                            h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_slice_sums: %s", hipGetErrorString(e));
    return CV_OK;
}

// ---- cv_slice_sums_noalloc: the slice sums with a static scratch array ------------------
// The partial sums of k_slice_sums (at most 1024 blocks x CV_NOALLOC_NZ slices x CV_SLICE_Q
// doubles, 10.5 MB) live in this file-scope device array, one copy per GPU for the life of
// the process, so the entry point allocates nothing.  Host code must not pass the array's
// name to a kernel (it is a host-side stand-in): its device address, per device, comes from
// hipGetSymbolAddress with that device current, taken once per device and kept below.
#define CV_NOALLOC_NZ 128
#define CV_NOALLOC_SCRATCH ((size_t)1024 * CV_NOALLOC_NZ * CV_SLICE_Q)
static __device__ __align__(256) double scratch[CV_NOALLOC_SCRATCH];
#define CV_MAX_DEVICES 64
static double *g_scratch_ptr[CV_MAX_DEVICES];
static std::mutex g_scratch_mu;

// device address of `scratch` on the current device h->device (set up on the first call)
static int scratch_ptr(int device, double **out) {
    if (device < 0 || device >= CV_MAX_DEVICES) return fail(CV_ERR_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(CV_ERR_HIP, "cv_slice_sums_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", device,
                        hipGetErrorString(e));
        g_scratch_ptr[device] = (double *)ptr;
    }
    *out = g_scratch_ptr[device];
    return CV_OK;
}

extern "C" int cv_slice_sums_noalloc(cv_handle *h, int which, double *d_out) {
    if (!h || !d_out) return fail(CV_ERR_ARG, "null argument");
    if (which < 0 || which > 2) return fail(CV_ERR_ARG, "which must be 0, 1 or 2");
    if (h->prm.nz > CV_NOALLOC_NZ)
        return fail(CV_ERR_CAPACITY, "cv_slice_sums_noalloc: %d eta slices exceed the %d of its static scratch array "
                                     "(use cv_slice_sums)", h->prm.nz, CV_NOALLOC_NZ);
    HIP_TRY(hipSetDevice(h->device));
    double *part = nullptr;
    int rc = scratch_ptr(h->device, &part);
    if (rc) return rc;
    launch_slice_sums(h, which, part, d_out);
    HIP_TRY(hipGetLastError());
    return CV_OK;
}

static int check_glauber(const cv_glauber *g) {
    if (!(g->num_nucleons >= 1.0f) || !(g->ra > 0.0f) || !(g->eta > 0.0f) || !(g->ro0 > 0.0f) ||
        !(g->si0 > 0.0f) || !(g->edmax >= 0.0f) || !(g->eta_gw > 0.0f))
        return fail(CV_ERR_ARG, "cv_glauber: NumOfNucleons >= 1 and positive Ra, Eta, NucleonDensity, Si0, Eta_gw "
                                "and Edmax >= 0 required");
    return CV_OK;
}

extern "C" int cv_glauber_ini(cv_handle *h, const cv_glauber *g, int which) {
    if (!h || !g) return fail(CV_ERR_ARG, "null argument");
    if (which < 0 || which > 2) return fail(CV_ERR_ARG, "which must be 0, 1 or 2");
    int rc = check_glauber(g);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(h->device));
    const int64_t nxy = (int64_t)h->prm.nx * h->prm.ny;
    float *edxy = nullptr;
    HIP_TRY(hipMalloc((void **)&edxy, sizeof(float) * nxy));
    hipLaunchKernelGGL(k_glauber_columns, dim3((unsigned)((nxy + 255) / 256)), dim3(256), 0, h->stream, h->phys, *g,
                       edxy);
    hipLaunchKernelGGL(k_glauber_ini, grid1d(h->n), dim3(256), 0, h->stream, h->phys, *g, edxy, h->ev[which]);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(edxy);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_glauber_ini: %s", hipGetErrorString(e));
    return CV_OK;
}

extern "C" int cv_glauber_nbinary(cv_handle *h, const cv_glauber *g, float xlow, float ylow, float dx, float dy,
                                  int nx, int ny, float *out) {
    if (!h || !g || !out || nx < 1 || ny < 1 || (int64_t)nx * ny > (1 << 28))
        return fail(CV_ERR_ARG, "cv_glauber_nbinary: bad argument");
    int rc = check_glauber(g);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(h->device));
    float *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(float) * (size_t)nx * ny));
    hipLaunchKernelGGL(k_glauber_nbinary, dim3((unsigned)(((int64_t)nx * ny + 255) / 256)), dim3(256), 0, h->stream,
                       *g, xlow, ylow, dx, dy, nx, ny, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, sizeof(float) * (size_t)nx * ny, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_glauber_nbinary: %s", hipGetErrorString(e));
    return CV_OK;
}

extern "C" int cv_smear_partons(cv_handle *h, const float *p4x4, int64_t npartons, float sigr, float sigz,
                                float kfactor, int which) {
    if (!h || (!p4x4 && npartons > 0) || npartons < 0) return fail(CV_ERR_ARG, "cv_smear_partons: bad argument");
    if (which < 0 || which > 2) return fail(CV_ERR_ARG, "which must be 0, 1 or 2");
    if (!(sigr > 0.0f) || !(sigz > 0.0f)) return fail(CV_ERR_ARG, "cv_smear_partons: SIGR, SIGZ must be positive");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    // the reference's constants (kernel_gaussian_smearing_new.cl:3-7) in double, rounded once
    const double c = 1.0 / (2.0 * (double)sigr * (double)sigr);
    const double cz = 1.0 / (2.0 * (double)sigz * (double)sigz);
    const double pif = (double)3.14159265358979323846f;      // M_PI_F
    const double w = (c / pif) * (sqrt(cz / pif) / (double)(float)h->prm.tau0);   // TAU0 is a float define
    float *dp = nullptr;
    SmearParton *ds = nullptr;
    const int64_t np1 = npartons > 0 ? npartons : 1;
    HIP_TRY(hipMalloc((void **)&dp, sizeof(float) * 8 * np1));
    hipError_t e = hipMalloc((void **)&ds, sizeof(SmearParton) * np1);
    if (e == hipSuccess && npartons > 0) e = hipMemcpy(dp, p4x4, sizeof(float) * 8 * npartons, hipMemcpyHostToDevice);
    if (e == hipSuccess && npartons > 0) {
        hipLaunchKernelGGL(k_smear_prep, dim3((unsigned)((npartons + 255) / 256)), dim3(256), 0, h->stream, dp,
                           npartons, ds);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        const cv_params &pp = h->prm;
        dim3 g((unsigned)((pp.nx + CV_SMEAR_BI - 1) / CV_SMEAR_BI), (unsigned)((pp.ny + CV_SMEAR_BJ - 1) / CV_SMEAR_BJ),
               (unsigned)((pp.nz + CV_SMEAR_BK - 1) / CV_SMEAR_BK));
        hipLaunchKernelGGL(k_smear, g, dim3(256), 0, h->stream, h->phys, h->eos, ds, npartons, (float)c,
                           (float)w, kfactor, h->ev[which]);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dp);
    (void)hipFree(ds);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_smear_partons: %s", hipGetErrorString(e));
    return CV_OK;
}

extern "C" int cv_longitudinal_profile(cv_handle *h, int which, const float *profile) {
    if (!h || !profile) return fail(CV_ERR_ARG, "null argument");
    if (which < 0 || which > 2) return fail(CV_ERR_ARG, "which must be 0, 1 or 2");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const int64_t nxy = (int64_t)h->prm.nx * h->prm.ny;
    float *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(float) * (4 * nxy + h->prm.nz)));
    hipError_t e = hipMemcpy(d + 4 * nxy, profile, sizeof(float) * h->prm.nz, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_long_profile_mid, dim3((unsigned)((nxy + 255) / 256)), dim3(256), 0, h->stream, h->phys,
                           h->ev[which], d);
        hipLaunchKernelGGL(k_long_profile, grid1d(h->n), dim3(256), 0, h->stream, h->phys, d, d + 4 * nxy,
                           h->ev[which]);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_longitudinal_profile: %s", hipGetErrorString(e));
    return CV_OK;
}

extern "C" int cv_sd_to_ed(cv_handle *h, int which, const double *s_nodes, const double *e_nodes, int64_t m) {
    if (!h || !s_nodes || !e_nodes || m < 2) return fail(CV_ERR_ARG, "cv_sd_to_ed: bad argument (2 nodes at least)");
    if (which < 0 || which > 2) return fail(CV_ERR_ARG, "which must be 0, 1 or 2");
    for (int64_t i = 1; i < m; i++)
        if (!(s_nodes[i] > s_nodes[i - 1]))
            return fail(CV_ERR_ARG, "cv_sd_to_ed: s nodes must increase strictly (node %lld)", (long long)i);
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    double *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(double) * 2 * m));
    hipError_t e = hipMemcpy(d, s_nodes, sizeof(double) * m, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + m, e_nodes, sizeof(double) * m, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_sd_to_ed, grid1d(h->n), dim3(256), 0, h->stream, h->n, h->ev[which], d, d + m, m);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_sd_to_ed: %s", hipGetErrorString(e));
    return CV_OK;
}

extern "C" int cv_polarization(int device, int kind, const float *sf, const float *om1, const float *om2,
                               int64_t nsf, const float *mom, int64_t nmom, float tfrz, float mass,
                               float *pol_lab, float *density, float *pol_lrf) {
    if (kind < 0 || kind > 2) return fail(CV_ERR_ARG, "cv_polarization: kind must be 0, 1 or 2");
    if (!sf || !om1 || (kind == 1 && !om2) || !mom || !pol_lab || !density || !pol_lrf || nsf < 0 || nmom < 1 ||
        nmom > (1 << 30))
        return fail(CV_ERR_ARG, "cv_polarization: bad argument");
    if (!(tfrz > 0.0f) || !(mass > 0.0f)) return fail(CV_ERR_ARG, "cv_polarization: T_frz and mass must be positive");
    HIP_TRY(hipSetDevice(device));
    const int w1 = kind == 1 ? 16 : 6;
    const int64_t n1 = nsf > 0 ? nsf : 1;
    size_t fl = (size_t)n1 * (8 + w1 + (kind == 1 ? 4 : 0)) + (size_t)nmom * 13;
    float *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(float) * fl));
    float *dsf = d, *d1 = dsf + 8 * n1, *d2 = d1 + (size_t)w1 * n1, *dm = d2 + (kind == 1 ? 4 * n1 : 0);
    float *dpl = dm + 4 * nmom, *dde = dpl + 4 * nmom, *dlr = dde + nmom;
    hipError_t e = hipSuccess;
    if (nsf > 0) {
        e = hipMemcpy(dsf, sf, sizeof(float) * 8 * nsf, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d1, om1, sizeof(float) * w1 * nsf, hipMemcpyHostToDevice);
        if (e == hipSuccess && kind == 1) e = hipMemcpy(d2, om2, sizeof(float) * 4 * nsf, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemcpy(dm, mom, sizeof(float) * 4 * nmom, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        PolP P;
        P.nsf = nsf; P.tfrz = tfrz; P.beta = (float)(1.0 / (double)tfrz); P.mass = mass; P.kind = kind;
        if (kind == 0)
            hipLaunchKernelGGL(k_polarization<0>, dim3((unsigned)nmom), dim3(256), 0, 0, P, dsf, d1, d2, dm, dpl, dde, dlr);
        else if (kind == 1)
            hipLaunchKernelGGL(k_polarization<1>, dim3((unsigned)nmom), dim3(256), 0, 0, P, dsf, d1, d2, dm, dpl, dde, dlr);
        else
            hipLaunchKernelGGL(k_polarization<2>, dim3((unsigned)nmom), dim3(256), 0, 0, P, dsf, d1, d2, dm, dpl, dde, dlr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(pol_lab, dpl, sizeof(float) * 4 * nmom, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(density, dde, sizeof(float) * nmom, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(pol_lrf, dlr, sizeof(float) * 4 * nmom, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_polarization: %s", hipGetErrorString(e));
    return CV_OK;
}

extern "C" int cv_sync(cv_handle *h) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CV_OK;
}

extern "C" int cv_last_elapsed_ms(cv_handle *h, float *ms) {
    if (!h || !ms) return fail(CV_ERR_ARG, "null argument");
    if (!h->timed) return fail(CV_ERR_STATE, "no timed call yet");
    HIP_TRY(hipEventSynchronize(h->ev_b));
    HIP_TRY(hipEventElapsedTime(ms, h->ev_a, h->ev_b));
    return CV_OK;
}

// ---- freeze-out hypersurface ------------------------------------------------
static void sf_free_out(cv_handle *h) {
    (void)hipFree(h->sf_out); (void)hipFree(h->sf_txyz); (void)hipFree(h->sf_pisf); (void)hipFree(h->sf_bksf);
    (void)hipFree(h->sf_om);
    h->sf_out = h->sf_txyz = h->sf_pisf = h->sf_bksf = h->sf_om = nullptr;
    h->sf_cap = 0;
}

extern "C" int cv_hypersf_reserve(cv_handle *h, int64_t capacity) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    if (capacity < 1 || capacity > ((int64_t)1 << 31))
        return fail(CV_ERR_ARG, "cv_hypersf_reserve: capacity %lld out of [1, 2^31]", (long long)capacity);
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    sf_free_out(h);
    bool ok = hipMalloc((void **)&h->sf_out, sizeof(float) * 8 * capacity) == hipSuccess &&
              hipMalloc((void **)&h->sf_txyz, sizeof(float) * 4 * capacity) == hipSuccess;
    if (ok && h->viscous) ok = hipMalloc((void **)&h->sf_pisf, sizeof(float) * 10 * capacity) == hipSuccess;
    if (ok && h->bulk) ok = hipMalloc((void **)&h->sf_bksf, sizeof(float) * capacity) == hipSuccess;
    if (ok && h->om[0]) ok = hipMalloc((void **)&h->sf_om, sizeof(float) * 32 * capacity) == hipSuccess;
    if (ok && !h->sf_count) {
        ok = hipMalloc((void **)&h->sf_count, sizeof(unsigned long long)) == hipSuccess &&
             hipMemset(h->sf_count, 0, sizeof(unsigned long long)) == hipSuccess;
    }
    if (!ok) {
        sf_free_out(h);
        return fail(CV_ERR_HIP, "cv_hypersf_reserve: allocation of %lld elements failed", (long long)capacity);
    }
    h->sf_cap = capacity;
    return CV_OK;
}

extern "C" int cv_hypersf_reset(cv_handle *h) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    HIP_TRY(hipSetDevice(h->device));
    if (h->sf_count) {
        HIP_TRY(hipMemsetAsync(h->sf_count, 0, sizeof(unsigned long long), h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    h->sf_have_old = false;
    return CV_OK;
}

extern "C" int cv_hypersf_snapshot(cv_handle *h, float tau) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    HIP_TRY(hipSetDevice(h->device));
    if (!h->sf_cap) {
        int rc = cv_hypersf_reserve(h, 1500000);
        if (rc) return rc;
    }
    if (!h->sf_ev) {
        bool ok = hipMalloc((void **)&h->sf_ev, sizeof(float) * 4 * h->n) == hipSuccess;
        if (ok && h->viscous) ok = hipMalloc((void **)&h->sf_pi, sizeof(float) * 10 * h->n) == hipSuccess;
        if (ok && h->bulk) ok = hipMalloc((void **)&h->sf_bk, sizeof(float) * h->n) == hipSuccess;
        if (!ok) return fail(CV_ERR_HIP, "cv_hypersf_snapshot: allocation of the previous-time fields failed");
    }
    HIP_TRY(hipMemcpyAsync(h->sf_ev, h->ev[1], sizeof(float) * 4 * h->n, hipMemcpyDeviceToDevice, h->stream));
    if (h->viscous)
        HIP_TRY(hipMemcpyAsync(h->sf_pi, h->pi[1], sizeof(float) * 10 * h->n, hipMemcpyDeviceToDevice, h->stream));
    if (h->bulk)
        HIP_TRY(hipMemcpyAsync(h->sf_bk, h->bk[1], sizeof(float) * h->n, hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->sf_tau_old = tau;
    h->sf_have_old = true;
    return CV_OK;
}

extern "C" int cv_hypersf_find(cv_handle *h, float tau_new) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    if (!h->sf_have_old) return fail(CV_ERR_STATE, "cv_hypersf_find: no snapshot (call cv_hypersf_snapshot first)");
    const cv_params &p = h->prm;
    if (p.nxskip < 1 || p.nyskip < 1 || p.nzskip < 1) return fail(CV_ERR_ARG, "cv_hypersf_find: skips must be >= 1");
    HIP_TRY(hipSetDevice(h->device));
    SfP s;
    s.nx = p.nx; s.ny = p.ny; s.nz = p.nz;
    s.nxskip = p.nxskip; s.nyskip = p.nyskip; s.nzskip = p.nzskip;
    s.gx = (p.nx - 1) / p.nxskip + 1; s.gy = (p.ny - 1) / p.nyskip + 1; s.gz = (p.nz - 1) / p.nzskip + 1;
    s.n = h->n;
    s.dx = p.dx; s.dy = p.dy; s.dz = p.dz; s.efrz = p.efrz;
    s.cap = h->sf_cap;
    s.viscous = h->viscous;
    int64_t total = (int64_t)s.gx * s.gy * s.gz;
    if (total >= ((int64_t)1 << 32)) return fail(CV_ERR_CAPACITY, "cv_hypersf_find: %lld cubes", (long long)total);
    if (total > h->sf_list_cap) {
        HIP_TRY(hipStreamSynchronize(h->stream));
        (void)hipFree(h->sf_list);
        h->sf_list = nullptr;
        h->sf_list_cap = 0;
        if (hipMalloc((void **)&h->sf_list, sizeof(uint32_t) * total) != hipSuccess)
            return fail(CV_ERR_HIP, "cv_hypersf_find: allocation of the cube list failed");
        h->sf_list_cap = total;
    }
    if (!h->sf_nlist && hipMalloc((void **)&h->sf_nlist, sizeof(unsigned)) != hipSuccess) {
        h->sf_nlist = nullptr;
        return fail(CV_ERR_HIP, "cv_hypersf_find: allocation failed");
    }
    // pass 1: list the crossing cubes; pass 2: their elements, grid-stride over the list
    hipLaunchKernelGGL(k_fix_reset, dim3(1), dim3(64), 0, h->stream, h->sf_nlist);
    const bool fp64 = (p.flags & CV_FLAG_SF_FP64) != 0;
    auto mark = fp64 ? k_hypersf_mark<double> : k_hypersf_mark<float>;
    hipLaunchKernelGGL(mark, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, h->stream, s, h->sf_ev, h->ev[1],
                       h->sf_list, h->sf_nlist);
    auto kern = fp64 ? k_hypersf<double> : k_hypersf<float>;
    const int64_t g2 = std::min<int64_t>((total + 127) / 128, (int64_t)(h->cus > 0 ? h->cus : 256) * 16);
    hipLaunchKernelGGL(kern, dim3((unsigned)g2), dim3(128), 0, h->stream, s, h->sf_list, h->sf_nlist, h->sf_ev,
                       h->ev[1], h->sf_pi, h->pi[1], h->bulk ? h->sf_bk : nullptr, h->bulk ? h->bk[1] : nullptr,
                       h->sf_tau_old, tau_new, h->sf_count, h->sf_out, h->sf_txyz, h->sf_pisf, h->sf_bksf,
                       h->om[0], h->om[1], h->sf_om);
    HIP_TRY(hipGetLastError());
    // d_omega[0] <- d_omega[1] after the search (pyvisc/visc.py:392-395)
    if (h->om[0])
        HIP_TRY(hipMemcpyAsync(h->om[0], h->om[1], sizeof(float) * CV_VORT_N * h->n, hipMemcpyDeviceToDevice,
                               h->stream));
    return cv_hypersf_snapshot(h, tau_new);
}

extern "C" int cv_vorticity_enable(cv_handle *h) {
    if (!h || !h->viscous) return fail(CV_ERR_STATE, "cv_vorticity_enable: viscous handle required");
    if (h->om[0]) return CV_OK;
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const size_t bytes = sizeof(float) * CV_VORT_N * (size_t)h->n;
    bool ok = hipMalloc((void **)&h->om[0], bytes) == hipSuccess && hipMalloc((void **)&h->om[1], bytes) == hipSuccess &&
              hipMemset(h->om[0], 0, bytes) == hipSuccess && hipMemset(h->om[1], 0, bytes) == hipSuccess;
    if (ok && h->sf_cap) ok = hipMalloc((void **)&h->sf_om, sizeof(float) * 32 * h->sf_cap) == hipSuccess;
    if (!ok) {
        (void)hipFree(h->om[0]); (void)hipFree(h->om[1]);
        h->om[0] = h->om[1] = nullptr;
        return fail(CV_ERR_HIP, "cv_vorticity_enable: allocation failed");
    }
    return CV_OK;
}

extern "C" int cv_vorticity(cv_handle *h, float tau) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    if (!h->om[0]) return fail(CV_ERR_STATE, "cv_vorticity: call cv_vorticity_enable first");
    if (!(tau > 0.0f)) return fail(CV_ERR_ARG, "cv_vorticity: tau must be positive");
    HIP_TRY(hipSetDevice(h->device));
    if (h->eos.kind == 1)
        hipLaunchKernelGGL(k_vorticity<1>, grid1d(h->n), dim3(256), 0, h->stream, h->phys, h->eos, h->ev[0],
                           h->ev[1], tau, h->om[1]);
    else
        hipLaunchKernelGGL(k_vorticity<0>, grid1d(h->n), dim3(256), 0, h->stream, h->phys, h->eos, h->ev[0],
                           h->ev[1], tau, h->om[1]);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CV_OK;
}

extern "C" int cv_vorticity_mu(cv_handle *h, float efrz, float *host, int64_t nfloats) {
    if (!h || !host) return fail(CV_ERR_ARG, "null argument");
    if (!h->om[0]) return fail(CV_ERR_STATE, "cv_vorticity_mu: call cv_vorticity_enable first");
    if (nfloats != 4 * h->n) return fail(CV_ERR_ARG, "cv_vorticity_mu: %lld floats expected", (long long)(4 * h->n));
    HIP_TRY(hipSetDevice(h->device));
    float *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(float) * 4 * h->n));
    if (h->eos.kind == 1)
        hipLaunchKernelGGL(k_vorticity_mu<1>, grid1d(h->n), dim3(256), 0, h->stream, h->phys, h->eos, h->ev[1],
                           h->om[1], efrz, d);
    else
        hipLaunchKernelGGL(k_vorticity_mu<0>, grid1d(h->n), dim3(256), 0, h->stream, h->phys, h->eos, h->ev[1],
                           h->om[1], efrz, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, d, sizeof(float) * 4 * h->n, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_vorticity_mu: %s", hipGetErrorString(e));
    return CV_OK;
}

extern "C" int cv_hypersf_get_vorticity(cv_handle *h, float *om, int64_t nelem) {
    if (!h || !om) return fail(CV_ERR_ARG, "null argument");
    if (!h->sf_om) return fail(CV_ERR_STATE, "cv_hypersf_get_vorticity: vorticity is not enabled");
    int64_t c = 0, cap = 0;
    int rc = cv_hypersf_count(h, &c, &cap);
    if (rc) return rc;
    if (c > cap) return fail(CV_ERR_CAPACITY, "cv_hypersf_get_vorticity: %lld elements, capacity %lld", (long long)c,
                             (long long)cap);
    if (nelem != c) return fail(CV_ERR_ARG, "cv_hypersf_get_vorticity: %lld elements stored", (long long)c);
    if (c) HIP_TRY(hipMemcpy(om, h->sf_om, sizeof(float) * 32 * c, hipMemcpyDeviceToHost));
    return CV_OK;
}

extern "C" int cv_hypersf_count(cv_handle *h, int64_t *count, int64_t *capacity) {
    if (!h || !count) return fail(CV_ERR_ARG, "null argument");
    unsigned long long c = 0;
    if (h->sf_count) {
        HIP_TRY(hipSetDevice(h->device));
        HIP_TRY(hipMemcpyAsync(&c, h->sf_count, sizeof(c), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    *count = (int64_t)c;
    if (capacity) *capacity = h->sf_cap;
    return CV_OK;
}

extern "C" int cv_hypersf_get(cv_handle *h, float *sf, float *txyz, float *pisf, float *bksf, int64_t nelem) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    int64_t c = 0;
    int rc = cv_hypersf_count(h, &c, nullptr);
    if (rc) return rc;
    if (c > h->sf_cap)
        return fail(CV_ERR_CAPACITY, "cv_hypersf_get: %lld elements found, capacity %lld (raise it with "
                    "cv_hypersf_reserve)", (long long)c, (long long)h->sf_cap);
    if (nelem != c) return fail(CV_ERR_ARG, "cv_hypersf_get: %lld elements stored, buffers for %lld",
                                (long long)c, (long long)nelem);
    if (c == 0) return CV_OK;
    if (sf) HIP_TRY(hipMemcpy(sf, h->sf_out, sizeof(float) * 8 * c, hipMemcpyDeviceToHost));
    if (txyz) HIP_TRY(hipMemcpy(txyz, h->sf_txyz, sizeof(float) * 4 * c, hipMemcpyDeviceToHost));
    if (pisf) {
        if (!h->sf_pisf) return fail(CV_ERR_STATE, "cv_hypersf_get: pi on the surface needs a viscous handle");
        HIP_TRY(hipMemcpy(pisf, h->sf_pisf, sizeof(float) * 10 * c, hipMemcpyDeviceToHost));
    }
    if (bksf) {
        if (!h->sf_bksf) return fail(CV_ERR_STATE, "cv_hypersf_get: bulk on the surface needs CV_FLAG_BULK");
        HIP_TRY(hipMemcpy(bksf, h->sf_bksf, sizeof(float) * c, hipMemcpyDeviceToHost));
    }
    return CV_OK;
}

// ---- event batches -----------------------------------------------------------
#include "cv_batch.h"

struct cv_batch {
    int device = 0;
    int bulk = 0;
    int ek = 1;
    bool fp64_root = false;                // CV_FLAG_FP64_ROOT: the SPEC stage kernels
    std::vector<cv_handle *> ev;           // one single-event state per event
    std::vector<int32_t> active;
    hipStream_t stream = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    bool timed = false;
    BatchState *d_state = nullptr;
    BatchEvent *d_events = nullptr;
    unsigned long long *d_ctr = nullptr;   // CV_BATCH_XCD_MAX work counters
    int nxcd = 8;                          // work ranges of the persistent stage (cv_batch_set_xcd_map)
    int64_t loop = 0;                      // loop index of the next step
    bool dirty = true;                     // event table / graph must be rebuilt
    int nact = 0;
    int64_t nwork = 0;
    int grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    cv_mlp *seed = nullptr;                // seeded batches: the stages run split (cv_batch_set_root_seed)
};

static void batch_drop_graph(cv_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

extern "C" int cv_batch_create(const cv_params *base, const cv_eos *eos, int nevents, const int32_t *dims,
                               int device, cv_batch **out) {
    if (!base || !eos || !dims || !out || nevents < 1) return fail(CV_ERR_ARG, "cv_batch_create: bad argument");
    cv_batch *b = new cv_batch();
    b->device = device;
    b->bulk = (base->flags & CV_FLAG_BULK) ? 1 : 0;
    if (base->flags & (CV_FLAG_RIEMANN_TEST | CV_FLAG_GUBSER_VISC_TEST | CV_FLAG_PIMN_OMEGA | CV_FLAG_OMEGA_OMEGA)) {
        delete b;
        return fail(CV_ERR_UNSUPPORTED, "cv_batch_create: the test flags and the vorticity couplings are single-event only");
    }
    b->fp64_root = (base->flags & CV_FLAG_FP64_ROOT) != 0;
    for (int e = 0; e < nevents; e++) {
        cv_params p = *base;
        p.nx = dims[3 * e]; p.ny = dims[3 * e + 1]; p.nz = dims[3 * e + 2];
        cv_handle *h = nullptr;
        int rc = cv_create(&p, eos, device, 1, &h);
        if (rc == CV_OK && h->kernel != 2)
            rc = fail(CV_ERR_UNSUPPORTED, "event %d: the batch needs the x-march kernel (grid below 2^32 / 40 cells, "
                                          "octave EOS of at most %d nodes)", e, CV_PTAB_LDS);
        if (rc != CV_OK) {
            if (h) cv_destroy(h);
            for (cv_handle *g : b->ev) cv_destroy(g);
            delete b;
            return rc;
        }
        b->ev.push_back(h);
    }
    b->ek = b->ev[0]->eos.kind;
    b->active.assign(nevents, 1);
    HIP_TRY(hipSetDevice(device));
    bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&b->ev_a) == hipSuccess && hipEventCreate(&b->ev_b) == hipSuccess &&
              hipMalloc((void **)&b->d_state, sizeof(BatchState)) == hipSuccess &&
              hipMalloc((void **)&b->d_events, sizeof(BatchEvent) * nevents) == hipSuccess &&
              hipMalloc((void **)&b->d_ctr, sizeof(unsigned long long) * CV_BATCH_XCD_MAX) == hipSuccess;
    if (!ok) {
        int rc = fail(CV_ERR_HIP, "cv_batch_create: allocation failed");
        cv_batch_destroy(b);
        return rc;
    }
    int cus = b->ev[0]->cus > 0 ? b->ev[0]->cus : 256;
    b->grid = cus * 3;
    *out = b;
    return CV_OK;
}

extern "C" int cv_batch_destroy(cv_batch *b) {
    if (!b) return CV_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    for (cv_handle *h : b->ev) cv_destroy(h);
    (void)hipFree(b->d_state); (void)hipFree(b->d_events); (void)hipFree(b->d_ctr);
    if (b->ev_a) (void)hipEventDestroy(b->ev_a);
    if (b->ev_b) (void)hipEventDestroy(b->ev_b);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CV_OK;
}

static int batch_event(cv_batch *b, int e, cv_handle **h) {
    if (!b) return fail(CV_ERR_ARG, "null batch");
    if (e < 0 || e >= (int)b->ev.size()) return fail(CV_ERR_ARG, "event %d out of [0, %d)", e, (int)b->ev.size());
    *h = b->ev[e];
    return CV_OK;
}

// the device descriptor must match the host-side slot pointers: bring the handles up to
// date (rotation of the steps run by the graph) before any per-event access
static int batch_settle(cv_batch *b) {
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return CV_OK;
}

extern "C" int cv_batch_handle(cv_batch *b, int event, cv_handle **out) {
    cv_handle *h;
    int rc = batch_event(b, event, &h);
    if (rc) return rc;
    rc = batch_settle(b);
    if (rc) return rc;
    if (!out) return fail(CV_ERR_ARG, "null argument");
    *out = h;
    b->dirty = true;               // the caller may change the tile: rebuild
    return CV_OK;
}

extern "C" int cv_batch_set_active(cv_batch *b, const int32_t *active) {
    if (!b || !active) return fail(CV_ERR_ARG, "null argument");
    int rc = batch_settle(b);
    if (rc) return rc;
    for (size_t e = 0; e < b->ev.size(); e++) b->active[e] = active[e] ? 1 : 0;
    b->dirty = true;
    return CV_OK;
}

// event table of the active events (largest work first) and the graph of one step
static int batch_build(cv_batch *b) {
    std::vector<int> order;
    for (size_t e = 0; e < b->ev.size(); e++)
        if (b->active[e]) order.push_back((int)e);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        return b->ev[a]->tile.nwork * b->ev[a]->tile.xc > b->ev[c]->tile.nwork * b->ev[c]->tile.xc;
    });
    // march chunk of the events whose chunk was not set (cv_set_march_chunk): one length for
    // the batch, so that the whole list holds about 32 work items per CU.  A lone event's
    // default fills the GPU by itself (1-plane chunks for a boost-invariant grid of two tile
    // columns); in a batch the other events bring the parallelism and longer chunks save the
    // chunk starts.  At most 16 planes, so that the list keeps enough items to balance.
    int64_t colplanes = 0;
    for (int e : order) {
        const cv_handle *h = b->ev[e];
        colplanes += (int64_t)h->tile.ntj * h->tile.ntk * h->prm.nx;
    }
    const int64_t target = (int64_t)(b->ev[0]->cus > 0 ? b->ev[0]->cus : 256) * 32;
    const int xcb = (int)std::max<int64_t>(1, std::min<int64_t>(16, colplanes / std::max<int64_t>(1, target)));
    std::vector<BatchEvent> tab(order.size());
    int64_t w = 0;
    for (size_t q = 0; q < order.size(); q++) {
        cv_handle *h = b->ev[order[q]];
        BatchEvent &t = tab[q];
        t.p = h->phys;
        t.tp = h->tile;
        // the batch has its own work ranges over the whole list: plain chunks along x
        t.tp.slabs = 0; t.tp.cps = 0; t.tp.xt = 0;
        if (h->march_xc == 0) t.tp.xc = std::min(xcb, h->prm.nx);
        t.tp.nchunks = (h->prm.nx + t.tp.xc - 1) / t.tp.xc;
        t.tp.nwork = (int64_t)t.tp.ntj * t.tp.ntk * t.tp.nchunks;
        t.work0 = w;
        for (int k = 0; k < 3; k++) { t.ev[k] = h->ev[k]; t.pi[k] = h->pi[k]; t.bk[k] = h->bk[k]; }
        t.Src = h->Src; t.IS_src = h->IS_src; t.bsrc = h->bsrc; t.udiff = h->udiff; t.good = h->good;
        t.fix = h->d_fix; t.fix_ctr = h->d_fix_ctr;
        t.cons = b->seed ? h->seed_buf : nullptr;
        t.seed = b->seed ? h->seed_buf + 9 * h->n : nullptr;
        w += t.tp.nwork;
    }
    if (w >= ((int64_t)1 << 31)) return fail(CV_ERR_CAPACITY, "batch: %lld work items exceed 2^31", (long long)w);
    BatchState st;
    st.tau0 = b->ev[0]->prm.tau0;
    st.dt = b->ev[0]->prm.dt;
    st.loop0 = b->loop;
    st.step = 0;
    st.nwork = w;
    st.nevents = (int)order.size();
    st.nxcd = b->nxcd > 1 ? b->nxcd : 1;
    HIP_TRY(hipSetDevice(b->device));
    if (!tab.empty())
        HIP_TRY(hipMemcpyAsync(b->d_events, tab.data(), sizeof(BatchEvent) * tab.size(), hipMemcpyHostToDevice,
                               b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_state, &st, sizeof(st), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->nact = (int)order.size();
    b->nwork = w;
    b->dirty = false;
    // one RK2 step: counter reset, stage 1, counter reset, stage 2, step + 1
    batch_drop_graph(b);
    if (b->nact == 0) return CV_OK;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    dim3 g(b->grid), blk(CV_TILE_THREADS);
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, b->stream, b->d_ctr);
    // seeded with an octave table: the network inside the stage (each wave seeds its cells
    // before the root solve; FP64 cells deferred as in the plain stages); with a row/col
    // table: split stages (the root solve's input to each event's cons), then the seed
    // network and the root solve over all events, FP64 cells solved there
    const bool seeded = b->seed != nullptr;
    const bool seed_in_stage = seeded && b->ek == 1;
    const bool seed_split = seeded && !seed_in_stage;
    const float *mP = seed_in_stage ? b->seed->P : nullptr;
    const uint32_t *mpk = seed_in_stage ? b->seed->pk : nullptr;
    const bool defer = b->fp64_root && !seed_split;
#define CV_BATCH_LAUNCH2(B, S, K, X) \
    hipLaunchKernelGGL((k_batch_stage<B, S, K, X>), g, blk, 0, b->stream, b->d_state, b->d_events, b->ev[0]->eos, \
                       b->d_ctr, mP, mpk)
#define CV_BATCH_LAUNCH(B, S)                                                              \
    do {                                                                                   \
        if (seed_in_stage) { if (b->fp64_root) CV_BATCH_LAUNCH2(B, S, 1, CV_SPEC_SEED | CV_SPEC_FP64); \
                             else CV_BATCH_LAUNCH2(B, S, 1, CV_SPEC_SEED); }                          \
        else if (seed_split) CV_BATCH_LAUNCH2(B, S, 0, CV_SPEC_SPLIT);                                \
        else if (b->ek == 1) { if (b->fp64_root) CV_BATCH_LAUNCH2(B, S, 1, CV_SPEC_FP64); else CV_BATCH_LAUNCH2(B, S, 1, 0); } \
        else { if (b->fp64_root) CV_BATCH_LAUNCH2(B, S, 0, CV_SPEC_FP64); else CV_BATCH_LAUNCH2(B, S, 0, 0); }            \
    } while (0)
    // deferred FP64 root solves of each stage, per event
    const int nact = b->nact;
    dim3 fg(std::max(1, 4 * (b->ev[0]->cus > 0 ? b->ev[0]->cus : 256) / std::max(1, nact)), nact), fb(256);
#define CV_BFIX_LAUNCH(B, S, K) hipLaunchKernelGGL((k_batch_fp64_fix<B, S, K>), fg, fb, 0, b->stream, b->d_state, \
                                                   b->d_events, b->ev[0]->eos)
    // seeded: the network over every event's cells, then the root solve (grids sized by the
    // largest event; blocks beyond an event's cells return at once)
    int64_t nmax = 0;
    for (int e = 0; e < (int)b->ev.size(); e++)
        if (b->active[e]) nmax = std::max<int64_t>(nmax, b->ev[e]->n);
    const int cus = b->ev[0]->cus > 0 ? b->ev[0]->cus : 256;
    dim3 sg((unsigned)std::min<int64_t>((nmax + MLP_SEED_PER_BLOCK - 1) / MLP_SEED_PER_BLOCK,
                                        std::max<int64_t>(1, 8 * cus / nact)), nact), sb(64 * MLP_WAVES);
    dim3 rg((unsigned)std::min<int64_t>((nmax + 255) / 256, std::max<int64_t>(1, 8 * cus / nact)), nact), rb(256);
#define CV_BROOT_LAUNCH(B, S, K) hipLaunchKernelGGL((k_batch_visc_root<B, S, K>), rg, rb, 0, b->stream, b->d_state, \
                                                    b->d_events, b->ev[0]->eos)
#define CV_BATCH_FIX(B, S)                                                                                    \
    do {                                                                                                      \
        if (seed_split) {                                                                                     \
            hipLaunchKernelGGL(k_batch_seed, sg, sb, 0, b->stream, b->d_state, b->d_events, b->seed->P,       \
                               b->seed->pk);                                                                  \
            CV_BROOT_LAUNCH(B, S, 0);                                                                         \
        } else if (defer) { if (b->ek == 1) CV_BFIX_LAUNCH(B, S, 1); else CV_BFIX_LAUNCH(B, S, 0); }          \
    } while (0)
    if (seeded) mlp_pack(b->seed, b->stream);
    if (defer)
        hipLaunchKernelGGL(k_batch_fix_reset, dim3((nact + 63) / 64), dim3(64), 0, b->stream, b->d_state, b->d_events);
    if (b->bulk) { CV_BATCH_LAUNCH(true, 1); CV_BATCH_FIX(true, 1); }
    else { CV_BATCH_LAUNCH(false, 1); CV_BATCH_FIX(false, 1); }
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, b->stream, b->d_ctr);
    if (defer)
        hipLaunchKernelGGL(k_batch_fix_reset, dim3((nact + 63) / 64), dim3(64), 0, b->stream, b->d_state, b->d_events);
    if (b->bulk) { CV_BATCH_LAUNCH(true, 2); CV_BATCH_FIX(true, 2); }
    else { CV_BATCH_LAUNCH(false, 2); CV_BATCH_FIX(false, 2); }
#undef CV_BROOT_LAUNCH
#undef CV_BATCH_FIX
#undef CV_BFIX_LAUNCH
#undef CV_BATCH_LAUNCH
#undef CV_BATCH_LAUNCH2
    hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(64), 0, b->stream, b->d_state);
    hipError_t e1 = hipStreamEndCapture(b->stream, &b->graph);
    if (e1 != hipSuccess) return fail(CV_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e1));
    HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return CV_OK;
}

extern "C" int cv_batch_steps_async(cv_batch *b, int64_t nsteps) {
    if (!b) return fail(CV_ERR_ARG, "null batch");
    if (nsteps < 0) return fail(CV_ERR_ARG, "negative step count");
    // the batched stage kernel carries no jet source: refuse rather than drop a jet set
    // on a member handle (cv_set_jet_source through cv_batch_handle)
    for (size_t e = 0; e < b->ev.size(); e++)
        if (b->active[e] && b->ev[e]->jet.on)
            return fail(CV_ERR_UNSUPPORTED, "cv_batch_steps: event %d has a jet source; jets are single-event only",
                        (int)e);
    HIP_TRY(hipSetDevice(b->device));
    for (cv_handle *h : b->ev) HIP_TRY(hipStreamSynchronize(h->stream));
    if (b->dirty) {
        int rc = batch_build(b);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(b->ev_a, b->stream));
    if (b->nact > 0)
        for (int64_t s = 0; s < nsteps; s++) HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    HIP_TRY(hipEventRecord(b->ev_b, b->stream));
    b->timed = true;
    // host view of the rotation: odd step counts swap slots 0 and 1 of active events
    if (nsteps & 1) {
        for (size_t e = 0; e < b->ev.size(); e++) {
            if (!b->active[e]) continue;
            cv_handle *h = b->ev[e];
            std::swap(h->ev[0], h->ev[1]); std::swap(h->pi[0], h->pi[1]); std::swap(h->bk[0], h->bk[1]);
        }
    }
    // the device step counter advanced with the host loop, and its buffer parity with the
    // host slot swaps above: the same graph serves the next call
    b->loop += nsteps;
    return CV_OK;
}

extern "C" int cv_batch_set_root_seed(cv_batch *b, cv_mlp *m) {
    if (!b) return fail(CV_ERR_ARG, "null batch");
    int rc = batch_settle(b);
    if (rc) return rc;
    // every event's own handle takes the network too (split stages, the same buffers), so
    // that stepping an event alone gives what the batch gives
    for (cv_handle *h : b->ev) {
        rc = m ? cv_set_root_seed(h, m, 0) : CV_OK;
        if (!m && h->seed_mlp) {
            h->seed_mlp = nullptr;
            h->root_split = false;
        }
        if (rc) return rc;
    }
    b->seed = m;
    b->dirty = true;
    return CV_OK;
}

extern "C" int cv_batch_set_xcd_map(cv_batch *b, int nxcd) {
    if (!b) return fail(CV_ERR_ARG, "null batch");
    if (nxcd < 0 || nxcd > CV_BATCH_XCD_MAX)
        return fail(CV_ERR_ARG, "cv_batch_set_xcd_map: nxcd in [0, %d]", CV_BATCH_XCD_MAX);
    int rc = batch_settle(b);
    if (rc) return rc;
    b->nxcd = nxcd;
    b->dirty = true;
    return CV_OK;
}

extern "C" int cv_batch_set_loop(cv_batch *b, int64_t loop) {
    if (!b) return fail(CV_ERR_ARG, "null batch");
    if (loop < 0) return fail(CV_ERR_ARG, "negative loop index");
    int rc = batch_settle(b);
    if (rc) return rc;
    b->loop = loop;
    b->dirty = true;
    return CV_OK;
}

extern "C" int cv_batch_sync(cv_batch *b) {
    if (!b) return fail(CV_ERR_ARG, "null batch");
    return batch_settle(b);
}

extern "C" int cv_batch_last_elapsed_ms(cv_batch *b, float *ms) {
    if (!b || !ms) return fail(CV_ERR_ARG, "null argument");
    if (!b->timed) return fail(CV_ERR_STATE, "no timed call yet");
    HIP_TRY(hipEventSynchronize(b->ev_b));
    HIP_TRY(hipEventElapsedTime(ms, b->ev_a, b->ev_b));
    return CV_OK;
}

extern "C" int cv_batch_info(cv_batch *b, int32_t *nevents, int64_t *loop, int64_t *nwork) {
    if (!b) return fail(CV_ERR_ARG, "null batch");
    if (nevents) *nevents = (int32_t)b->ev.size();
    if (loop) *loop = b->loop;
    if (nwork) {
        // the list of the last build; before the first build, the events' own chunks
        int64_t w = 0;
        for (size_t e = 0; e < b->ev.size(); e++)
            if (b->active[e]) w += b->ev[e]->tile.nwork;
        *nwork = b->dirty ? w : b->nwork;
    }
    return CV_OK;
}

// ---- learned root-solve seed -------------------------------------------------------

static uint64_t mlp_rng(uint64_t &s) {             // SplitMix64
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

extern "C" int cv_mlp_nparam(void) { return MLP_NPARAM; }

extern "C" int cv_mlp_create(int device, uint64_t seed, cv_mlp **out) {
    if (!out) return fail(CV_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(device));
    cv_mlp *m = new cv_mlp();
    m->device = device;
    bool ok = hipMalloc((void **)&m->P, sizeof(float) * MLP_NPARAM) == hipSuccess &&
              hipMalloc((void **)&m->m1, sizeof(float) * MLP_NPARAM) == hipSuccess &&
              hipMalloc((void **)&m->m2, sizeof(float) * MLP_NPARAM) == hipSuccess &&
              hipMalloc((void **)&m->grad, sizeof(float) * MLP_NPARAM) == hipSuccess &&
              hipMalloc((void **)&m->loss, sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&m->pk, sizeof(uint32_t) * MLP_PK_WORDS) == hipSuccess &&
              hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess;
    if (!ok) {
        cv_mlp_destroy(m);
        return fail(CV_ERR_HIP, "cv_mlp_create: allocation failed");
    }
    // uniform(-a, a) with a = sqrt(6 / (fan_in + fan_out)); zero biases
    std::vector<float> h(MLP_NPARAM, 0.0f);
    uint64_t s = seed;
    auto fill = [&](int off, int fo, int fi) {
        double a = sqrt(6.0 / (fi + fo));
        for (int q = 0; q < fo * fi; q++) h[off + q] = (float)((2.0 * ((mlp_rng(s) >> 11) * (1.0 / 9007199254740992.0)) - 1.0) * a);
    };
    fill(MLP_OFF_W1, MLP_H, MLP_IN);
    fill(MLP_OFF_W2, MLP_H, MLP_H);
    fill(MLP_OFF_W3, MLP_OUT, MLP_H);
    std::vector<float> z(MLP_NPARAM, 0.0f);
    HIP_TRY(hipMemcpy(m->P, h.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->m1, z.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->m2, z.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    *out = m;
    return CV_OK;
}

extern "C" int cv_mlp_destroy(cv_mlp *m) {
    if (!m) return CV_OK;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    (void)hipFree(m->P); (void)hipFree(m->m1); (void)hipFree(m->m2); (void)hipFree(m->grad); (void)hipFree(m->loss);
    (void)hipFree(m->part); (void)hipFree(m->part_loss); (void)hipFree(m->X); (void)hipFree(m->T);
    (void)hipFree(m->pk);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return CV_OK;
}

extern "C" int cv_mlp_get_params(cv_mlp *m, float *host, int n) {
    if (!m || !host || n != MLP_NPARAM) return fail(CV_ERR_ARG, "cv_mlp_get_params: need %d floats", MLP_NPARAM);
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(host, m->P, sizeof(float) * n, hipMemcpyDeviceToHost));
    return CV_OK;
}

extern "C" int cv_mlp_set_params(cv_mlp *m, const float *host, int n) {
    if (!m || !host || n != MLP_NPARAM) return fail(CV_ERR_ARG, "cv_mlp_set_params: need %d floats", MLP_NPARAM);
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(m->P, host, sizeof(float) * n, hipMemcpyHostToDevice));
    m->adam_t = 0;
    HIP_TRY(hipMemset(m->m1, 0, sizeof(float) * MLP_NPARAM));
    HIP_TRY(hipMemset(m->m2, 0, sizeof(float) * MLP_NPARAM));
    return CV_OK;
}

static int mlp_parts(cv_mlp *m, int64_t nblk) {
    if (m->part_blocks >= nblk) return CV_OK;
    (void)hipFree(m->part); (void)hipFree(m->part_loss);
    m->part = nullptr; m->part_loss = nullptr; m->part_blocks = 0;
    HIP_TRY(hipMalloc((void **)&m->part, sizeof(float) * MLP_NPARAM * nblk));
    HIP_TRY(hipMalloc((void **)&m->part_loss, sizeof(float) * nblk));
    m->part_blocks = nblk;
    return CV_OK;
}

static const int64_t MLP_SAMPLES_PER_BLOCK = 16 * MLP_WAVES;

// gradient of the loss on device arrays X, T (n samples) into m->grad, loss into m->loss;
// adam: also one Adam step
static int mlp_grad_dev(cv_mlp *m, const float *X, const float *T, int64_t n, int adam, float lr) {
    const int64_t nblk = (n + MLP_SAMPLES_PER_BLOCK - 1) / MLP_SAMPLES_PER_BLOCK;
    int rc = mlp_parts(m, nblk);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)nblk), dim3(64 * MLP_WAVES), 0, m->stream, m->P, X, T, n,
                       1.0f / (float)n, m->part, m->part_loss);
    HIP_TRY(hipGetLastError());
    float c1 = 1.0f, c2 = 1.0f;
    const float b1 = 0.9f, b2 = 0.999f;
    if (adam) {
        m->adam_t++;
        c1 = (float)(1.0 / (1.0 - pow((double)b1, (double)m->adam_t)));
        c2 = (float)(1.0 / (1.0 - pow((double)b2, (double)m->adam_t)));
    }
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, m->part,
                       m->part_loss, (int)nblk, m->grad, m->loss, m->P, m->m1, m->m2, lr, b1, b2, 1e-8f, c1, c2, adam);
    HIP_TRY(hipGetLastError());
    return CV_OK;
}

extern "C" int cv_mlp_grad(cv_mlp *m, const float *X, const float *T, int64_t n, float *grad, float *loss) {
    if (!m || !X || !T || !grad || !loss || n < 1) return fail(CV_ERR_ARG, "cv_mlp_grad: bad argument");
    HIP_TRY(hipSetDevice(m->device));
    float *dX = nullptr, *dT = nullptr;
    HIP_TRY(hipMalloc((void **)&dX, sizeof(float) * MLP_IN * n));
    hipError_t e = hipMalloc((void **)&dT, sizeof(float) * MLP_OUT * n);
    if (e == hipSuccess) e = hipMemcpy(dX, X, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dT, T, sizeof(float) * MLP_OUT * n, hipMemcpyHostToDevice);
    int rc = e == hipSuccess ? mlp_grad_dev(m, dX, dT, n, 0, 0.0f) : fail(CV_ERR_HIP, "%s", hipGetErrorString(e));
    if (!rc) {
        e = hipStreamSynchronize(m->stream);
        if (e == hipSuccess) e = hipMemcpy(grad, m->grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(CV_ERR_HIP, "cv_mlp_grad: %s", hipGetErrorString(e));
    }
    (void)hipFree(dX); (void)hipFree(dT);
    return rc;
}

extern "C" int cv_mlp_set_data(cv_mlp *m, const float *X, const float *T, int64_t n) {
    if (!m || !X || !T || n < 1) return fail(CV_ERR_ARG, "cv_mlp_set_data: bad argument");
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipStreamSynchronize(m->stream));
    (void)hipFree(m->X); (void)hipFree(m->T);
    m->X = m->T = nullptr; m->ndata = 0;
    HIP_TRY(hipMalloc((void **)&m->X, sizeof(float) * MLP_IN * n));
    HIP_TRY(hipMalloc((void **)&m->T, sizeof(float) * MLP_OUT * n));
    HIP_TRY(hipMemcpy(m->X, X, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->T, T, sizeof(float) * MLP_OUT * n, hipMemcpyHostToDevice));
    m->ndata = n;
    return CV_OK;
}

// steps Adam steps on consecutive batches of the training set (wrapping around); the loss
// of every step goes to loss_hist (may be NULL)
extern "C" int cv_mlp_train(cv_mlp *m, int steps, int batch, float lr, float *loss_hist) {
    if (!m || steps < 0 || batch < 1) return fail(CV_ERR_ARG, "cv_mlp_train: bad argument");
    if (m->ndata < batch) return fail(CV_ERR_STATE, "cv_mlp_train: training set smaller than a batch");
    HIP_TRY(hipSetDevice(m->device));
    std::vector<float> hist(steps > 0 ? steps : 1);
    float *d_hist = nullptr;
    if (loss_hist && steps > 0) HIP_TRY(hipMalloc((void **)&d_hist, sizeof(float) * steps));
    int64_t off = 0;
    int rc = CV_OK;
    for (int s = 0; s < steps && !rc; s++) {
        if (off + batch > m->ndata) off = 0;
        rc = mlp_grad_dev(m, m->X + off * MLP_IN, m->T + off * MLP_OUT, batch, 1, lr);
        if (!rc && d_hist) {
            hipError_t e = hipMemcpyAsync(d_hist + s, m->loss, sizeof(float), hipMemcpyDeviceToDevice, m->stream);
            if (e != hipSuccess) rc = fail(CV_ERR_HIP, "%s", hipGetErrorString(e));
        }
        off += batch;
    }
    hipError_t e = hipStreamSynchronize(m->stream);
    if (!rc && e != hipSuccess) rc = fail(CV_ERR_HIP, "cv_mlp_train: %s", hipGetErrorString(e));
    if (!rc && d_hist) {
        e = hipMemcpy(loss_hist, d_hist, sizeof(float) * steps, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(CV_ERR_HIP, "%s", hipGetErrorString(e));
    }
    (void)hipFree(d_hist);
    return rc;
}

extern "C" int cv_mlp_forward(cv_mlp *m, const float *X, int64_t n, float *Y) {
    if (!m || !X || !Y || n < 1) return fail(CV_ERR_ARG, "cv_mlp_forward: bad argument");
    HIP_TRY(hipSetDevice(m->device));
    float *dX = nullptr, *dY = nullptr;
    HIP_TRY(hipMalloc((void **)&dX, sizeof(float) * MLP_IN * n));
    hipError_t e = hipMalloc((void **)&dY, sizeof(float) * MLP_OUT * n);
    if (e == hipSuccess) e = hipMemcpy(dX, X, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const int64_t nblk = (n + MLP_SAMPLES_PER_BLOCK - 1) / MLP_SAMPLES_PER_BLOCK;
        hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)nblk), dim3(64 * MLP_WAVES), 0, m->stream, m->P, dX, n, dY);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e == hipSuccess) e = hipMemcpy(Y, dY, sizeof(float) * MLP_OUT * n, hipMemcpyDeviceToHost);
    (void)hipFree(dX); (void)hipFree(dY);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_mlp_forward: %s", hipGetErrorString(e));
    return CV_OK;
}

extern "C" int cv_mlp_features(cv_mlp *m, const float *T00, const float *M, const float *T03, int64_t n, float *X) {
    if (!m || !T00 || !M || !T03 || !X || n < 1) return fail(CV_ERR_ARG, "cv_mlp_features: bad argument");
    HIP_TRY(hipSetDevice(m->device));
    float *buf = nullptr;
    HIP_TRY(hipMalloc((void **)&buf, sizeof(float) * n * (3 + MLP_IN)));
    float *dT = buf, *dM = buf + n, *d3 = buf + 2 * n, *dX = buf + 3 * n;
    hipError_t e = hipMemcpy(dT, T00, sizeof(float) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dM, M, sizeof(float) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d3, T03, sizeof(float) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_seed_features, grid1d(n), dim3(256), 0, m->stream, dT, dM, d3, n, dX);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e == hipSuccess) e = hipMemcpy(X, dX, sizeof(float) * MLP_IN * n, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_mlp_features: %s", hipGetErrorString(e));
    return CV_OK;
}

// root solve of host states (T^tt, |T^ti|, T^t eta) with the handle's EOS: plain
// (m == NULL) or from the network's |v| seed; outputs e, |v|, iterations
extern "C" int cv_root_eval(cv_handle *h, cv_mlp *m, const float *T00, const float *M, const float *T03, int64_t n,
                            float *e_out, float *v_out, int32_t *iters) {
    if (!h || !T00 || !M || !T03 || !e_out || !v_out || !iters || n < 1) return fail(CV_ERR_ARG, "cv_root_eval: bad argument");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    float *buf = nullptr;
    int *dit = nullptr;
    // T00, M, T03, X (4), Y (2), e, v
    HIP_TRY(hipMalloc((void **)&buf, sizeof(float) * n * 11));
    hipError_t e = hipMalloc((void **)&dit, sizeof(int) * n);
    float *dT = buf, *dM = buf + n, *d3 = buf + 2 * n, *dX = buf + 3 * n, *dY = buf + 7 * n, *de = buf + 9 * n,
          *dv = buf + 10 * n;
    if (e == hipSuccess) e = hipMemcpy(dT, T00, sizeof(float) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dM, M, sizeof(float) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d3, T03, sizeof(float) * n, hipMemcpyHostToDevice);
    const float *seed = nullptr;
    if (e == hipSuccess && m) {
        hipLaunchKernelGGL(k_seed_features, grid1d(n), dim3(256), 0, h->stream, dT, dM, d3, n, dX);
        const int64_t nblk = (n + MLP_SAMPLES_PER_BLOCK - 1) / MLP_SAMPLES_PER_BLOCK;
        hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)nblk), dim3(64 * MLP_WAVES), 0, h->stream, m->P, dX, n, dY);
        // |v| is the second output of each pair: pass every other float via a strided copy
        hipLaunchKernelGGL(k_take_col, grid1d(n), dim3(256), 0, h->stream, dY, n, 2, 1, de);
        seed = de;
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        float *seed_buf = nullptr;
        if (seed) {           // keep the seed apart from the outputs
            e = hipMalloc((void **)&seed_buf, sizeof(float) * n);
            if (e == hipSuccess) e = hipMemcpyAsync(seed_buf, seed, sizeof(float) * n, hipMemcpyDeviceToDevice, h->stream);
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_root_eval, grid1d(n), dim3(256), 0, h->stream, h->eos, dT, dM, seed_buf, n, de, dv, dit);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        (void)hipFree(seed_buf);
    }
    if (e == hipSuccess) e = hipMemcpy(e_out, de, sizeof(float) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(v_out, dv, sizeof(float) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(iters, dit, sizeof(int) * n, hipMemcpyDeviceToHost);
    (void)hipFree(buf); (void)hipFree(dit);
    if (e != hipSuccess) return fail(CV_ERR_HIP, "cv_root_eval: %s", hipGetErrorString(e));
    return CV_OK;
}

// ideal stage with the learned seed: conserved state (dry run), features, network, seeded
// root solve (the stage proper)
static int launch_ideal_seeded(cv_handle *h, const StageP &st, float *ev_st, float *ev1, float *ev_out) {
    const int64_t n = h->n;
    float *cons = h->seed_buf, *X = cons + 3 * n, *Y = cons + 7 * n, *seed = cons + 9 * n;
    hipLaunchKernelGGL(k_ideal_stage, grid1d(n), dim3(256), 0, h->stream, h->phys, h->eos, st, ev_st, ev1, h->Src,
                       ev_out, cons, nullptr, nullptr, h->jet);
    hipLaunchKernelGGL(k_seed_features, grid1d(n), dim3(256), 0, h->stream, cons, cons + n, cons + 2 * n, n, X);
    const int64_t nblk = (n + MLP_SAMPLES_PER_BLOCK - 1) / MLP_SAMPLES_PER_BLOCK;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)nblk), dim3(64 * MLP_WAVES), 0, h->stream, h->seed_mlp->P, X, n, Y);
    hipLaunchKernelGGL(k_take_col, grid1d(n), dim3(256), 0, h->stream, Y, n, 2, 1, seed);
    hipLaunchKernelGGL(k_ideal_stage, grid1d(n), dim3(256), 0, h->stream, h->phys, h->eos, st, ev_st, ev1, h->Src,
                       ev_out, nullptr, seed, h->track_iters ? h->iters : nullptr, h->jet);
    HIP_TRY(hipGetLastError());
    return CV_OK;
}

extern "C" int cv_set_root_seed(cv_handle *h, cv_mlp *m, int track_iters) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    if (m && !h->viscous && (h->prm.flags & CV_FLAG_FP64_ROOT))
        return fail(CV_ERR_UNSUPPORTED, "cv_set_root_seed: the seeded ideal stage has no FP64 root solve");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (m && !h->seed_buf) HIP_TRY(hipMalloc((void **)&h->seed_buf, sizeof(float) * 10 * h->n));
    if (track_iters && !h->iters) {
        HIP_TRY(hipMalloc((void **)&h->iters, sizeof(int) * h->n));
        HIP_TRY(hipMemset(h->iters, 0, sizeof(int) * h->n));
    }
    // a network set on a viscous handle without one starts in the split form
    if (h->viscous && m && !h->seed_mlp) h->root_split = true;
    h->seed_mlp = m;
    h->track_iters = track_iters != 0;
    return CV_OK;
}

extern "C" int cv_set_root_split(cv_handle *h, int on) {
    if (!h) return fail(CV_ERR_ARG, "null handle");
    if (!h->viscous) return fail(CV_ERR_STATE, "cv_set_root_split: viscous handle required");
    if (on && h->kernel == 0) return fail(CV_ERR_UNSUPPORTED, "cv_set_root_split: needs a tiled stage kernel (kind 1 or 2)");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (on && !h->seed_buf) HIP_TRY(hipMalloc((void **)&h->seed_buf, sizeof(float) * 10 * h->n));
    if (on && h->jet.on) return fail(CV_ERR_UNSUPPORTED, "the jet source does not run with the split root solve");
    h->root_split = on != 0;
    return CV_OK;
}

extern "C" int cv_get_root_iters(cv_handle *h, int32_t *host, int64_t n) {
    if (!h || !host) return fail(CV_ERR_ARG, "null argument");
    if (!h->iters) return fail(CV_ERR_STATE, "iteration counts are not tracked (cv_set_root_seed)");
    if (n != h->n) return fail(CV_ERR_ARG, "cv_get_root_iters: %lld cells", (long long)h->n);
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(host, h->iters, sizeof(int) * n, hipMemcpyDeviceToHost));
    return CV_OK;
}
