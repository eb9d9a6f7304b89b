// cv_batch.h -- one persistent launch per RK stage for an ensemble of events with their
// own grids (event-by-event hydro, the reference's pyvisc/ebe.py running one CLVisc per
// event).  Every event keeps its own fields, tile and march chunk; the work items of all
// events (tile column x x-chunk) form one list, largest events first, that the workgroups
// of a persistent grid take from atomic counters: the list is cut into nxcd contiguous
// ranges with a counter each, workgroup b starts on range b % nxcd (the XCD the hardware
// dispatched it to, so that neighbouring tiles share that XCD's L2) and moves on to the
// following ranges once its own is exhausted.  The step index lives in device
// memory so that a captured graph of one step can be replayed: tau, the stage constants
// and the d_ev[0] / d_ev[1] buffer rotation follow from it, with the same FP32 operations
// as the single-event path.
#pragma once
#include "cv_tiled.h"
#include "cv_mlp.h"

namespace cv {

struct BatchEvent {
    PhysP p;
    TileP tp;
    int64_t work0;                     // first work item of this event in the list
    float *ev[3], *pi[3], *bk[3];      // physical buffers of slots 0..2 at step parity 0
    float *Src, *IS_src, *bsrc, *udiff, *good;
    FixEntry *fix;                     // deferred FP64 root solves (CV_FLAG_FP64_ROOT), n entries
    unsigned *fix_ctr;
    float *cons, *seed;                // seeded batches: the root solve's input (4 n) and the seeds (n)
};

struct BatchState {
    double tau0, dt;                   // TAU0, DT as the configuration holds them
    int64_t loop0;                     // loop index of step 0 of this descriptor
    int64_t step;                      // steps taken since the descriptor was built
    int64_t nwork;                     // work items of all active events
    int nevents;                       // active events (entries of the event table)
    int nxcd;                          // work ranges / counters, 1..CV_BATCH_XCD_MAX
};

#define CV_BATCH_XCD_MAX 16

__global__ void k_batch_reset(unsigned long long *ctr) {
    if (threadIdx.x < CV_BATCH_XCD_MAX && blockIdx.x == 0) ctr[threadIdx.x] = 0ull;
}

__global__ void k_batch_advance(BatchState *bs) {
    if (threadIdx.x == 0 && blockIdx.x == 0) bs->step = bs->step + 1;
}

template <int EK>
__device__ __forceinline__ auto make_batch_eos(const EosP &E, const float *pt) {
    if constexpr (EK == 1) return EosOctP(E, E.tab, pt);
    else return EosT<EK>(E, E.tab);
}

// SPEC: 0, CV_SPEC_FP64 (the FP64 root-solve branch for fast cells, CV_FLAG_FP64_ROOT),
// CV_SPEC_SEED (seeded batches with an octave table: the network in the stage, mlp_P and
// mlp_pk its parameters and packed operands), CV_SPEC_SEED | CV_SPEC_FP64, or CV_SPEC_SPLIT
// (seeded batches with a row/col table: the stage writes the root solve's input to each
// event's cons and k_batch_seed / k_batch_visc_root finish the cells).
template <bool BULK, int STEP, int EK, int SPEC>
__global__ __launch_bounds__(CV_TILE_THREADS, 3) void k_batch_stage(const BatchState *__restrict__ bs,
                                                                     const BatchEvent *__restrict__ evt, EosP E,
                                                                     unsigned long long *ctr,
                                                                     const float *__restrict__ mlp_P,
                                                                     const uint32_t *__restrict__ mlp_pk) {
    __shared__ float s_c[CF_N][CV_TILE_MAXC];
    __shared__ float s_f[15][CV_TILE_MAXF];
    __shared__ int s_work;
    // octave table: its pressure column in LDS, as in the single-event marching kernel
    __shared__ float s_pt[EK == 1 ? CV_PTAB_LDS : 1];
    if constexpr (EK == 1)
        for (int k = threadIdx.x; k <= E.nlast; k += CV_TILE_THREADS) s_pt[k] = E.tab[k].y;
    const auto eos = make_batch_eos<EK>(E, s_pt);
    const int64_t n = bs->step;
    const int par = (int)(n & 1);
    // tau of this stage as cv_visc_steps_async computes it on the host (tau_at)
    const float tau = (float)(bs->tau0 + (double)(bs->loop0 + n + (STEP - 1)) * bs->dt);
    const int64_t nwork = bs->nwork;
    const int nev = bs->nevents;
    const int nr = bs->nxcd;
    const int64_t per = nwork / nr, rem = nwork % nr;
    const int home = (int)(blockIdx.x % (unsigned)nr);
    int k = 0;                                          // thread 0: ranges exhausted so far
    for (;;) {
        if (threadIdx.x == 0) {
            int64_t got = nwork;
            for (; k < nr; ++k) {
                const int x = (home + k) % nr;
                const int64_t cnt = per + (x < rem ? 1 : 0);
                if (cnt == 0) continue;
                const unsigned long long t = atomicAdd(&ctr[x], 1ull);
                if (t < (unsigned long long)cnt) {
                    got = (int64_t)x * per + (x < rem ? x : rem) + (int64_t)t;
                    break;
                }
            }
            s_work = (int)got;
        }
        __syncthreads();
        const int w = __builtin_amdgcn_readfirstlane(s_work);     // uniform: event table via scalar loads
        __syncthreads();
        if (w >= nwork) break;
        int lo = 0, hi = nev - 1;                       // last event with work0 <= w
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (evt[mid].work0 <= w) lo = mid; else hi = mid - 1;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        const BatchEvent &ev = evt[lo];
        StageP st;
        st.step = STEP;
        st.tau = tau;
        st.r_tau = 1.0f / tau;
        st.r_taudz = 1.0f / (tau * ev.p.dz);
        st.old_time = tau - (float)(STEP - 1) * ev.p.dt;
        st.r_step = 1.0f / (float)STEP;
        float *E0 = ev.ev[par], *E1 = ev.ev[1 - par], *E2 = ev.ev[2];
        float *P0 = ev.pi[par], *P1 = ev.pi[1 - par], *P2 = ev.pi[2];
        float *B0 = ev.bk[par], *B1 = ev.bk[1 - par], *B2 = ev.bk[2];
        StagePtrs q;
        float *cons = (SPEC & CV_SPEC_SPLIT) ? ev.cons : nullptr;
        if (STEP == 1) {
            q = StagePtrs{E1, P1, B1, E1, E2, P1, B1, ev.udiff, ev.IS_src, ev.bsrc, ev.Src, P2, B2, E2, ev.good,
                          nullptr, cons, ev.fix, ev.fix_ctr, mlp_P, mlp_pk, nullptr, nullptr};
        } else {
            q = StagePtrs{E2, P2, B2, E1, E2, P1, B1, ev.udiff, ev.IS_src, ev.bsrc, ev.Src, P0, B0, E0, ev.good,
                          ev.udiff, cons, ev.fix, ev.fix_ctr, mlp_P, mlp_pk, nullptr, nullptr};
        }
        tile_work<BULK, STEP, EK, SPEC, true>(ev.p, eos, st, ev.tp, q, w - ev.work0, s_c, s_f);
    }
}

// Deferred FP64 root solves of the active events (blockIdx.y = entry of the event table),
// with the slots of this step's parity as k_batch_stage uses them.
template <bool BULK, int STEP, int EK>
__global__ __launch_bounds__(256) void k_batch_fp64_fix(const BatchState *__restrict__ bs,
                                                        const BatchEvent *__restrict__ evt, EosP E) {
    if ((int)blockIdx.y >= bs->nevents) return;
    const BatchEvent &ev = evt[blockIdx.y];
    const int par = (int)(bs->step & 1);
    float *E0 = ev.ev[par], *E1 = ev.ev[1 - par], *E2 = ev.ev[2];
    float *out = STEP == 1 ? E2 : E0;
    const unsigned cnt = *ev.fix_ctr;
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gridDim.x * blockDim.x)
        fp64_fix_cell<BULK, STEP, EK>(ev.p, E, ev.fix[q], E1, out, ev.udiff);
}

__global__ void k_batch_fix_reset(const BatchState *__restrict__ bs, const BatchEvent *__restrict__ evt) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < bs->nevents) *evt[e].fix_ctr = 0u;
}

}  // namespace cv

namespace cv {

// Seed network over the root-solve input of every active event (blockIdx.y = entry of the
// event table): the seeds of the single-event k_mlp_seed_cons, sample by sample.
__global__ __launch_bounds__(64 * MLP_WAVES) void k_batch_seed(const BatchState *__restrict__ bs,
                                                             const BatchEvent *__restrict__ evt,
                                                             const float *__restrict__ P,
                                                             const uint32_t *__restrict__ pk) {
    if ((int)blockIdx.y >= bs->nevents) return;
    const BatchEvent &ev = evt[blockIdx.y];
    const int64_t n = ev.p.n;
    const int64_t nblk = (n + MLP_SEED_PER_BLOCK - 1) / MLP_SEED_PER_BLOCK;
    for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) seed_cons_block(P, pk, ev.cons, n, ev.seed, blk);
}

// The end of update_ev for every cell of every active event (blockIdx.y = entry of the event
// table), with the slots of this step's parity as k_batch_stage uses them: FP64 for fast
// cells (CV_FLAG_FP64_ROOT), else from the event's seeds.
template <bool BULK, int STEP, int EK>
__global__ __launch_bounds__(256) void k_batch_visc_root(const BatchState *__restrict__ bs,
                                                         const BatchEvent *__restrict__ evt, EosP E) {
    if ((int)blockIdx.y >= bs->nevents) return;
    const BatchEvent &ev = evt[blockIdx.y];
    const int par = (int)(bs->step & 1);
    float *E0 = ev.ev[par], *E1 = ev.ev[1 - par], *E2 = ev.ev[2];
    const float *Bout = STEP == 1 ? ev.bk[2] : ev.bk[par];
    float *out = STEP == 1 ? E2 : E0;
    const int64_t n = ev.p.n;
    for (int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; I < n; I += (int64_t)gridDim.x * blockDim.x)
        visc_root_cell<BULK, STEP, EK>(ev.p, E, I, ev.cons, Bout, E1, ev.seed, out, STEP == 2 ? ev.udiff : nullptr,
                                       nullptr);
}

}  // namespace cv
