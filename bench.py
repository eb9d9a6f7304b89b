#!/usr/bin/env python
"""Throughput of the viscous KT + Israel-Stewart (+ bulk) hydro step on one MI355X.

Workload: 385 x 385 x 115 cells, DX = DY = 0.08 fm, DZ = 0.15, DTAU = 0.02 fm,
TAU0 = 0.6 fm, eta/s = 0.08, bulk zeta/s(T) peaked at T = 0.155 GeV; initial energy
density = 400 Gaussian hotspots (width 0.5 fm) at Woods-Saxon nucleon positions
(R = 6.62, a = 0.546) times an eta plateau of half-width 1.3 with Gaussian tails
(sigma 1.5), peak 50 GeV/fm^3; smooth crossover EOS in the octave table.
A step is one full RK2 step over all cells; the unit of work is a cell-update.

Prints one JSON line (stdout).  Diagnostics go to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

NX, NY, NZ = 385, 385, 115
DX, DY, DZ, DT, TAU0 = 0.08, 0.08, 0.15, 0.02, 0.6
BULK = (0.05, 0.155, 0.02)          # zeta/s peak, T_peak [GeV], width [GeV]
HBM_PEAK_GBS = 8000.0
REF_CELL_UPDATES = 1.69e7           # reference README: 1.01 s/step at 385x385x115 (BSZ=64)

# Compulsory HBM bytes per cell of each launch of one step (FP32; cell = 1 cell-update).
# stage 1: read ev_n 16, pi_n 40, udiff 16, stale e 4;
#          write pi* 40, ev* 16, stage-1 sources IS_src 40 + Src 16, goodcell 4      -> 192
# stage 2: read ev* 16, pi* 40, ev_n 16, pi_n 40, stage-1 sources 56;
#          write pi_{n+1} 40, ev_{n+1} 16, udiff 16, goodcell 4                      -> 244
# bulk Pi adds 4 B per read/write: stage 1 +12 (Pi_n; Pi*, its source), stage 2 +16.
ALGO_BYTES = {False: [192, 244], True: [204, 260]}
LAUNCH_NAMES = ['visc_stage1', 'visc_stage2']


PROFILE = os.path.join(HERE, 'profiles', 'k_visc_tiled_visc_bulk_385x385x115.md')


def profiled_value(key, kernel):
    """A per-dispatch figure of a stage kernel from the committed rocprofv3 counter summary
    (lines 'key kernel value'), or None."""
    try:
        with open(PROFILE) as f:
            for line in f:
                w = line.split()
                if len(w) == 3 and w[0] == key and w[1] == kernel:
                    return float(w[2])
    except OSError:
        pass
    return None


def profiled_traffic(kernel):
    """HBM-side bytes per dispatch of a stage kernel (FETCH_SIZE x 2.00 + WRITE_SIZE,
    calibrated on a device copy), or None."""
    return profiled_value('traffic_bytes', kernel)


VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4     # wave64 VALU instructions per second: 1024 SIMDs, 4 cycles each


def log(*a):
    print(*a, file=sys.stderr)
    sys.stderr.flush()


def make_workload(bulk=True, ic=True):
    from clvisc_amd import _lib
    from clvisc_amd import ic as ic_mod
    from clvisc_amd.eos import Eos
    eos = Eos('crossover')
    tab = eos.create_table()
    p = _lib.CvParams()
    p.nx, p.ny, p.nz = NX, NY, NZ
    p.tau0, p.dt = TAU0, DT
    p.dx, p.dy, p.dz = DX, DY, DZ
    p.etaos_xmin, p.etaos_ymin, p.etaos_left_slop, p.etaos_right_slop, p.lam1 = 0.0, 0.08, 0.0, 0.0, -10.0
    p.flags = 0
    if bulk:
        p.flags |= _lib.CV_FLAG_BULK
        p.zetaos_max, p.zetaos_tpeak, p.zetaos_width = BULK
    p.efrz = float(eos.f_ed(0.137))
    p.nxskip = p.nyskip = p.nzskip = 1
    if not ic:
        return p, tab, None
    t0 = time.time()
    pos = ic_mod.woods_saxon_positions(400, 20240601)
    ev = ic_mod.hotspots(NX, NY, NZ, DX, DY, DZ, positions=pos, emax=50.0)
    log('IC built in %.1fs' % (time.time() - t0))
    return p, tab, ev


def host_info():
    """CPU model and counts of the host (lscpu), the CPUs this process may use and the
    OpenMP thread setting."""
    import subprocess
    info = {'cpus': os.cpu_count(), 'affinity_cpus': len(os.sched_getaffinity(0)),
            'OMP_NUM_THREADS': os.environ.get('OMP_NUM_THREADS')}
    keys = {'Model name': 'model', 'Socket(s)': 'sockets', 'Core(s) per socket': 'cores_per_socket',
            'Thread(s) per core': 'threads_per_core'}
    try:
        out = subprocess.run(['lscpu'], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(':')
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_threads_default():
    """The host's CPU share: OMP_NUM_THREADS when set (16 per GPU on the GPU hosts), else
    the CPUs this process may run on."""
    try:
        return max(1, int(os.environ.get('OMP_NUM_THREADS', '')))
    except ValueError:
        return len(os.sched_getaffinity(0))


def _oracle_visc_rate(p, tab, ev, nz, k0, nthreads, warmup, steps):
    """FP32 oracle viscous steps on the eta slices k0 .. k0+nz-1 of the benchmark field:
    (cell-updates/s over the timed steps, seconds)."""
    from oracle import oracle as O
    from clvisc_amd import _lib
    sub = ev.reshape(NX, NY, NZ, 4)[:, :, k0:k0 + nz, :].reshape(-1, 4)
    bulk = (p.zetaos_max, p.zetaos_tpeak, p.zetaos_width) if p.flags & _lib.CV_FLAG_BULK else None
    op = O.make_params(NX, NY, nz, DT, DX, DY, DZ, TAU0, tab, etaos_xmin=0.0, etaos_ymin=0.08,
                       bulk=bulk, nthreads=nthreads)
    o = O.Oracle(op, 'f32', viscous=True)
    o.load_ini(sub)
    o.visc_steps(warmup)
    t0 = time.perf_counter()
    o.visc_steps(steps)
    dt = time.perf_counter() - t0
    del o
    return NX * NY * nz * steps / dt, dt


def cpu_baseline(p, tab, ev, nthreads):
    """The FP32 oracle (unfused, reference-ordered kernels, OpenMP) on the host: the full
    385x385x115 grid of the same initial condition, 2 warm-up and 10 timed RK2 steps on the
    host's CPU share, and one thread on a 385x385x5 central eta slab (1 warm-up, 3 timed)."""
    rate, dt = _oracle_visc_rate(p, tab, ev, NZ, 0, nthreads, 2, 10)
    nz1 = 5
    rate1, dt1 = _oracle_visc_rate(p, tab, ev, nz1, NZ // 2 - nz1 // 2, 1, 1, 3)
    return {'value': rate, 'unit': 'cell-updates/s', 'cores': nthreads, 'kind': 'port',
            'ms_per_step': 1000.0 * dt / 10,
            'sample': '%dx%dx%d (the whole workload), 2 warm-up + 10 timed RK2 steps in %.1f s, FP32 oracle '
                      '(unfused reference-ordered kernels, OpenMP, %d threads)' % (NX, NY, NZ, dt, nthreads),
            'one_thread': {'value': rate1, 'unit': 'cell-updates/s', 'cores': 1,
                           'sample': '%dx%dx%d central eta slab, 1 warm-up + 3 timed RK2 steps in %.1f s'
                                     % (NX, NY, nz1, dt1)},
            'host': host_info(),
            'reference_published': {'value': 2.64e6, 'unit': 'cell-updates/s', 'ms_per_step': 6450.0,
                                    'note': 'reference README, viscous 385x385x115 on an OpenCL CPU device '
                                            '(BSZ=16), hardware not stated; shear only'}}


C1_N, C1_STEPS = 201, 200


def c1_gubser(nthreads, with_cpu=True):
    """Config 1: ideal Gubser flow (q = 1/fm, conformal ideal-gas EOS, e0 at T = 0.3 GeV,
    tau0 = 1 fm) on 201x201x1, DX = 0.08 fm, DT = 0.02 fm, 200 steps: GPU ms/step and
    cell-updates/s, the FP32 oracle on the host at nthreads and at 1 thread, and the L1 error
    of e against the analytic flow (r < 4 fm) at the end."""
    from clvisc_amd import _lib, ic as ic_mod
    from clvisc_amd.eos import Eos
    eos = Eos('ideal_gas')
    tab = eos.create_table()
    p = _lib.CvParams()
    p.nx, p.ny, p.nz = C1_N, C1_N, 1
    p.tau0, p.dt, p.dx, p.dy, p.dz = 1.0, 0.02, 0.08, 0.08, 0.15
    p.nxskip = p.nyskip = p.nzskip = 1
    e0 = float(eos.f_ed(0.3))
    ev, norm = ic_mod.gubser_ideal(C1_N, C1_N, 1, 0.08, 0.08, 0.15, 1.0, q=1.0, e0=e0)
    ev = ev.astype(np.float32)
    h = _lib.Handle(p, tab, viscous=False)
    h.set('ev1', ev)
    h.ideal_steps(0, 10)
    h.set('ev1', ev)
    h.sync()
    t0 = time.perf_counter()
    h.ideal_steps(0, C1_STEPS, sync=False)
    h.sync()
    wall = time.perf_counter() - t0
    g = h.get('ev1')
    h.close()
    tau_end = float(np.float32(1.0 + C1_STEPS * 0.02))
    ea, _ = ic_mod.gubser_ideal(C1_N, C1_N, 1, 0.08, 0.08, 0.15, tau_end, q=1.0)
    ea[:, 0] *= norm
    x = (np.arange(C1_N) - C1_N // 2) * 0.08
    X, Y = np.meshgrid(x, x, indexing='ij')
    m = np.sqrt(X * X + Y * Y).ravel() < 4.0
    l1 = float(np.abs(g[m, 0] - ea[m, 0]).sum() / ea[m, 0].sum())
    ncell = C1_N * C1_N
    out = {'workload': 'ideal Gubser flow 201x201x1, 200 RK2 steps, FP32', 'gpu_ms_per_step': 1000.0 * wall / C1_STEPS,
           'gpu_value': ncell * C1_STEPS / wall, 'unit': 'cell-updates/s', 'l1_e_vs_analytic': l1}
    if with_cpu:
        from oracle import oracle as O
        for key, nth in (('cpu', nthreads), ('cpu_one_thread', 1)):
            op = O.make_params(C1_N, C1_N, 1, 0.02, 0.08, 0.08, 0.15, 1.0, tab, nthreads=nth)
            o = O.Oracle(op, 'f32', viscous=False)
            o.load_ini(ev)
            t0 = time.perf_counter()
            o.ideal_steps(C1_STEPS)
            dt = time.perf_counter() - t0
            out[key] = {'ms_per_step': 1000.0 * dt / C1_STEPS, 'value': ncell * C1_STEPS / dt, 'threads': nth,
                        'bit_equal_gpu': bool(np.array_equal(o.ev[1], g))}
    return out


def ensemble_events():
    """64 seeded events: 56 boost-invariant N x N grids, N = 161 .. 301 (odd), and eight
    161 x 161 x 41 grids; 50 .. 400 hotspots each (event e: 50 + 350 e / 63), seed 1000 + e."""
    dims = []
    for i in range(56):
        n = 161 + 2 * int(round(i * 70 / 55.0))
        dims.append((n, n, 1))
    dims += [(161, 161, 41)] * 8
    nhot = [50 + int(round(350 * e / 63.0)) for e in range(64)]
    return dims, nhot


def ensemble_ic(d, nhot, seed):
    from clvisc_amd import ic
    pos = ic.woods_saxon_positions(nhot, seed)
    return ic.hotspots(d[0], d[1], d[2], DX, DY, DZ, positions=pos, emax=50.0)


def run_ensemble(args, p, tab, name):
    """Batched event-by-event workload: one persistent launch per RK stage over all
    events (graph-replayed steps) vs the same events stepped one after another.  With
    --learned also the learned + batched form: the same events in a second batch whose root
    solves are seeded by the network (inside the persistent stage launch), trained first on
    the split states of one 161x161x41 event of the ensemble; reported against both plain
    forms, with the largest relative e difference from the plain batch after the same steps."""
    from clvisc_amd import _lib
    from clvisc_amd.autotune import Autotuner
    dims, nhot = ensemble_events()
    t0 = time.time()
    evs = [ensemble_ic(d, nhot[e], 1000 + e) for e, d in enumerate(dims)]
    log('ensemble IC built in %.1fs: %d events, %d cells' % (time.time() - t0, len(dims),
                                                             sum(a * b * c for a, b, c in dims)))
    batch = _lib.Batch(p, tab, dims)
    for e in range(len(dims)):
        h = batch.handle(e)
        h.visc_initialize(TAU0)
        h.ideal_stage(1, np.float32(TAU0))
        h.set('ev1', evs[e])
    tuned = None
    if args.autotune:
        t0 = time.time()
        tuner = Autotuner(p, tab, path=args.autotune_cache or None)
        reps = {}
        for e, d in enumerate(dims):
            # one representative per family: 2D grids in steps of 20 cells, the 3D shape
            rep = d if d[2] > 1 else (min(301, 161 + 20 * int(round((d[0] - 161) / 20.0))),) * 2 + (1,)
            if rep not in reps:
                reps[rep] = tuner.tune(rep, ensemble_ic(rep, 200, 7))
            ty, tz, xc, _ = reps[rep]
            if d[2] == 1:
                nt = -(-rep[1] // ty)
                ty = -(-d[1] // nt)
            h = batch.handle(e)
            h.set_stage_kernel(2, ty, tz)
            h.set_march_chunk(xc)
        tuned = {'%dx%dx%d' % k: list(v[:3]) for k, v in reps.items()}
        log('autotuned %d shapes in %.1fs: %s' % (len(reps), time.time() - t0, tuned))
    ncell = sum(a * b * c for a, b, c in dims)
    batch.steps(args.warmup)
    batch.sync()
    t0 = time.perf_counter()
    batch.steps(args.steps, sync=False)
    batch.sync()
    wall = time.perf_counter() - t0
    value = ncell * args.steps / wall
    log('batched: %d steps %.3fs -> %.3f ms/step, %.4g cell-updates/s (device %.3f ms)'
        % (args.steps, wall, 1000 * wall / args.steps, value, batch.last_elapsed_ms()))
    edmax = max(batch.handle(e).max_ed(1) for e in range(len(dims)))
    # the same events, one launch sequence per event and step
    singles = []
    for e, d in enumerate(dims):
        q = _lib.CvParams.from_buffer_copy(p)
        q.nx, q.ny, q.nz = d
        hs = _lib.Handle(q, tab, viscous=True)
        hs.visc_initialize(TAU0)
        hs.ideal_stage(1, np.float32(TAU0))
        hs.set('ev1', evs[e])
        singles.append(hs)
    nseq = max(2, args.steps // 4)
    for loop in range(args.warmup):
        for hs in singles:
            hs.visc_steps(loop, 1, sync=False)
    for hs in singles:
        hs.sync()
    t0 = time.perf_counter()
    for loop in range(args.warmup, args.warmup + nseq):
        for hs in singles:
            hs.visc_steps(loop, 1, sync=False)
    for hs in singles:
        hs.sync()
    seq = ncell * nseq / (time.perf_counter() - t0)
    log('one event after another: %.4g cell-updates/s (%d steps)' % (seq, nseq))
    for hs in singles:
        hs.close()
    learned = None
    if args.learned:
        learned = ensemble_learned(args, p, tab, dims, evs, batch, value, seq)
    out = {
        'metric': 'cell-updates/s, batched event-by-event viscous KT+IS (64 uneven events)',
        'value': value, 'unit': 'cell-updates/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': 1000.0 * wall / args.steps,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': '64 events: 56 boost-invariant N^2 (N=161..301) + 8 x 161x161x41, 50-400 '
                               'Woods-Saxon hotspots each, viscous + bulk, crossover EOS; one persistent '
                               'launch per RK stage, graph-replayed steps',
                   'cells': ncell, 'device': name, 'autotuned': tuned,
                   'sequential_value': seq, 'batched_over_sequential': value / seq,
                   'physical': bool(np.isfinite(edmax) and 0.0 < edmax <= 50.0)},
    }
    if learned is not None:
        out['learned_batched'] = learned
    print(json.dumps(out))
    batch.close()


def ensemble_learned(args, p, tab, dims, evs, batch, plain_value, seq_value):
    """The learned + batched form of the ensemble (see run_ensemble); batch is the plain batch
    after its warm-up and timed steps."""
    from clvisc_amd import _lib, surrogate
    e3 = next(e for e, d in enumerate(dims) if d[2] > 1)
    q = _lib.CvParams.from_buffer_copy(p)
    q.nx, q.ny, q.nz = dims[e3]
    h = _lib.Handle(q, tab, viscous=True)
    h.visc_initialize(TAU0)
    h.ideal_stage(1, np.float32(TAU0))
    h.set('ev1', evs[e3])
    t0 = time.perf_counter()
    states = surrogate.visc_run_states(h, 0, 6, every=2)
    h.close()
    mlp, hist = surrogate.train_seed(states, steps=2000, batch=4096, lr=3e-3, nmax=5_000_000)
    t_train = time.perf_counter() - t0
    log('ensemble learned: seed trained on %d states of event %d (%s) in %.1fs, loss %.4g -> %.4g'
        % (states[0].size, e3, 'x'.join(map(str, dims[e3])), t_train, hist[:20].mean(), hist[-20:].mean()))
    del states
    bl = _lib.Batch(p, tab, dims)
    for e in range(len(dims)):
        hh = bl.handle(e)
        if args.autotune:                       # the plain batch's tuned tiles and chunks
            ref = batch.handle(e)
            hh.set_stage_kernel(*ref.stage_kernel())
            hh.set_march_chunk(ref.march_chunk()[0])
        hh.visc_initialize(TAU0)
        hh.ideal_stage(1, np.float32(TAU0))
        hh.set('ev1', evs[e])
    bl.set_root_seed(mlp)
    ncell = sum(a * b * c for a, b, c in dims)
    bl.steps(args.warmup)
    bl.sync()
    t0 = time.perf_counter()
    bl.steps(args.steps, sync=False)
    bl.sync()
    wall = time.perf_counter() - t0
    value = ncell * args.steps / wall
    rel = 0.0
    for e in range(len(dims)):
        ep, el = batch.handle(e).get('ev1')[:, 0], bl.handle(e).get('ev1')[:, 0]
        big = ep > 1e-3 * ep.max()
        rel = max(rel, float((np.abs(el[big] - ep[big]) / ep[big]).max()))
    log('learned + batched: %d steps %.3fs -> %.4g cell-updates/s: %.3fx the plain batch, %.3fx the events '
        'one after another; max rel e vs the plain batch %.3g' % (args.steps, wall, value, value / plain_value,
                                                                    value / seq_value, rel))
    bl.close()
    mlp.close()
    # FP32 sensitivity of the plain batch itself: the same steps from the initial fields with e
    # scaled by (1 + 2^-23) (one ulp), compared with the plain batch
    bu = _lib.Batch(p, tab, dims)
    for e in range(len(dims)):
        hh = bu.handle(e)
        if args.autotune:
            ref = batch.handle(e)
            hh.set_stage_kernel(*ref.stage_kernel())
            hh.set_march_chunk(ref.march_chunk()[0])
        hh.visc_initialize(TAU0)
        hh.ideal_stage(1, np.float32(TAU0))
        evp = np.array(evs[e], np.float32, copy=True)
        evp[:, 0] *= np.float32(1.0 + 2.0 ** -23)
        hh.set('ev1', evp)
    bu.steps(args.warmup + args.steps)
    rel_ulp = 0.0
    for e in range(len(dims)):
        ep, eu = batch.handle(e).get('ev1')[:, 0], bu.handle(e).get('ev1')[:, 0]
        big = ep > 1e-3 * ep.max()
        rel_ulp = max(rel_ulp, float((np.abs(eu[big] - ep[big]) / ep[big]).max()))
    bu.close()
    log('plain batch from the initial fields perturbed by one ulp: max rel e vs the plain batch %.3g' % rel_ulp)
    return {'value': value, 'unit': 'cell-updates/s', 'ms_per_step': 1000.0 * wall / args.steps,
            'over_plain_batched': value / plain_value, 'over_plain_sequential': value / seq_value,
            'max_rel_e_vs_plain_batched': rel, 'matched': bool(rel <= MATCHED_REL),
            'plain_one_ulp_ic_max_rel_e': rel_ulp, 'within_one_ulp_sensitivity': bool(rel <= rel_ulp),
            'train_seconds': t_train, 'loss_last': float(hist[-20:].mean())}


MATCHED_REL = 1e-4     # learned form counts as matched when max rel |e_learned - e_plain| <= this


def run_learned(args, p, tab, ev, h_plain, loop_end):
    """The learned form of the same workload: the root solve seeded by the 4-64-64-2 network
    trained on the split stage's own states of this run, with the network inside the stage
    kernel (fused: each wave evaluates it on its 64 cells before the root solve).  Same
    initial state and step count as the plain run that ended at loop_end on h_plain.  The
    split form (stage kernel, then the network and the root solve in their own kernels) is
    timed over a window of the following steps.  Returns the 'learned' object of the JSON
    line."""
    from clvisc_amd import _lib, surrogate
    ncell = NX * NY * NZ
    hl = _lib.Handle(p, tab, viscous=True)
    hl.visc_initialize(TAU0)
    hl.ideal_stage(1, np.float32(TAU0))
    hl.set('ev1', ev)
    w = args.warmup
    t0 = time.perf_counter()
    states = surrogate.visc_run_states(hl, 0, w, every=max(1, w // 2))
    t_states = time.perf_counter() - t0
    t0 = time.perf_counter()
    mlp, hist = surrogate.train_seed(states, steps=2000, batch=4096, lr=3e-3, nmax=5_000_000)
    t_train = time.perf_counter() - t0
    log('learned: %d states from %d split steps (%.1fs); trained 2000 x 4096 in %.2fs, loss %.4g -> %.4g'
        % (states[0].size, w, t_states, t_train, hist[:20].mean(), hist[-20:].mean()))
    del states
    hl.set_root_seed(mlp, track_iters=False)
    hl.set_root_split(False)
    nrun = loop_end - w
    hl.sync()
    t0 = time.perf_counter()
    hl.visc_steps(w, nrun, sync=False)
    hl.sync()
    wall = time.perf_counter() - t0
    value = ncell * nrun / wall
    # matched error: the fields of the two runs after the same steps from the same start
    e_p, e_l = h_plain.get('ev1')[:, 0], hl.get('ev1')[:, 0]
    big = e_p > 1e-3 * e_p.max()
    rel = float((np.abs(e_l[big] - e_p[big]) / e_p[big]).max())
    # iterations per cell of one more step each (tracked), at the same loop index
    hl.set_root_seed(mlp, track_iters=True)
    hl.visc_steps(loop_end, 1)
    it_l = hl.root_iters()
    # the split form of the learned step, over the next steps of the same run
    nsplit = max(2, min(20, nrun))
    hl.set_root_seed(mlp, track_iters=False)
    hl.set_root_split(True)
    hl.sync()
    t0 = time.perf_counter()
    hl.visc_steps(loop_end + 1, nsplit, sync=False)
    hl.sync()
    split_learned_value = ncell * nsplit / (time.perf_counter() - t0)
    h_plain.set_root_split(True)
    h_plain.set_root_seed(None, track_iters=True)
    h_plain.visc_steps(loop_end, 1)
    it_p = h_plain.root_iters()
    # the split form without the seed (the network's share of the learned step time)
    h_plain.set_root_seed(None, track_iters=False)
    h_plain.sync()
    t0 = time.perf_counter()
    h_plain.visc_steps(loop_end + 1, nsplit, sync=False)
    h_plain.sync()
    split_value = ncell * nsplit / (time.perf_counter() - t0)
    log('learned (network in the stage kernel): %d steps %.3fs -> %.3f ms/step, %.4g cell-updates/s; '
        'split form %.4g (without seed %.4g); iterations plain %.3f learned %.3f; max rel e vs plain %.3g'
        % (nrun, wall, 1000 * wall / nrun, value, split_learned_value, split_value, it_p.mean(), it_l.mean(), rel))
    # FP32 sensitivity of the plain run itself: the same steps from the initial field with
    # e scaled by (1 + 2^-23) (one ulp), compared with the plain run at loop_end
    hl.close()
    hu = _lib.Handle(p, tab, viscous=True)
    hu.visc_initialize(TAU0)
    hu.ideal_stage(1, np.float32(TAU0))
    evp = np.array(ev, np.float32, copy=True)
    evp[:, 0] *= np.float32(1.0 + 2.0 ** -23)
    hu.set('ev1', evp)
    del evp
    hu.visc_steps(0, loop_end)
    e_u = hu.get('ev1')[:, 0]
    rel_ulp = float((np.abs(e_u[big] - e_p[big]) / e_p[big]).max())
    log('plain run from the initial field perturbed by one ulp: max rel e vs plain %.3g after %d steps'
        % (rel_ulp, loop_end))
    hu.close()
    mlp.close()
    return {'value': value, 'unit': 'cell-updates/s', 'ms_per_step': 1000.0 * wall / nrun, 'steps': nrun,
            'form': 'fused (the network in the stage kernel)', 'split_value': split_learned_value,
            'max_rel_e_vs_plain': rel, 'matched': bool(rel <= MATCHED_REL), 'matched_tolerance': MATCHED_REL,
            'plain_one_ulp_ic_max_rel_e': rel_ulp,
            'within_one_ulp_sensitivity': bool(rel <= rel_ulp),
            'iterations_plain': float(it_p.mean()), 'iterations_learned': float(it_l.mean()),
            'split_unseeded_value': split_value, 'train_seconds': t_train,
            'loss_first': float(hist[:20].mean()), 'loss_last': float(hist[-20:].mean()),
            'learned_over_plain': None}


def combined_run(p, tab, nev, max_loops=3000, sf_capacity=6_000_000, sf_out='', check_event=None, seed_mlp=None):
    """`nev` batched 385x385x115 viscous + bulk events (400 hotspots each, their own seeds), FP64
    root solve for cells with |T^ti| > 0.99 T^tt, evolved until every event's max e drops below
    e(T = 0.137 GeV) (checked every ntskip = 20 steps, as CLVisc.evolve does), freeze-out surface
    found every ntskip steps (nxskip = nyskip = 3, nzskip = 1), FP64 energy balance of one step in
    every ntskip.  Finished events leave the batch.  The plain GPU form (the same events stepped
    one after another) is timed over a window of steps.  check_event: an event index whose
    first ntskip steps are also run on its own handle and compared bit for bit.  seed_mlp: the
    learned + batched form (cv_batch_set_root_seed: split stages, the network's seeds for the
    root solve, FP64 for the fast cells).  Returns a dict."""
    from clvisc_amd import _lib
    from clvisc_amd import ic as ic_mod
    ntskip = 20
    p.flags |= _lib.CV_FLAG_FP64_ROOT
    p.fp64_root_vmin = 0.99
    p.nxskip, p.nyskip, p.nzskip = 3, 3, 1
    efrz = p.efrz
    ncell = NX * NY * NZ
    dims = [(NX, NY, NZ)] * nev
    t_start = time.perf_counter()
    batch = _lib.Batch(p, tab, dims)
    hs = [batch.handle(e) for e in range(nev)]
    seeds = [20240601 + e for e in range(nev)]
    single = None
    for e, h in enumerate(hs):
        pos = ic_mod.woods_saxon_positions(400, seeds[e])
        ev = ic_mod.hotspots(NX, NY, NZ, DX, DY, DZ, positions=pos, emax=50.0)
        h.visc_initialize(TAU0)
        h.ideal_stage(1, np.float32(TAU0))
        h.set('ev1', ev)
        h.hypersf_reserve(sf_capacity)
        h.hypersf_snapshot(np.float32(TAU0))
        if check_event == e:
            single = _lib.Handle(p, tab, viscous=True)
            single.visc_initialize(TAU0)
            single.ideal_stage(1, np.float32(TAU0))
            single.set('ev1', ev)
    del ev
    if seed_mlp is not None:
        batch.set_root_seed(seed_mlp)
    log('combined: %d events of %dx%dx%d set up in %.1fs (e_frz %.4g GeV/fm^3)'
        % (nev, NX, NY, NZ, time.perf_counter() - t_start, efrz))
    tau_of = lambda n: float(np.float32(TAU0 + n * DT))
    active = [1] * nev
    steps_done = [0] * nev
    finish_tau = [None] * nev
    balance = []                       # (event, loop, relative residual)
    energy = [[] for _ in range(nev)]  # (tau, E) per event
    check = None
    loop = 0
    dev_ms = 0.0
    t_loop = time.perf_counter()
    t_last = t_loop
    while sum(active) and loop < max_loops:
        # one step with the FP64 energy balance of the active events around it
        pre = {e: hs[e].energy_sum(tau_of(loop), 1) for e in range(nev) if active[e]}
        batch.steps(1)
        dev_ms += batch.last_elapsed_ms()
        for e, (E0, W0) in pre.items():
            E1, _ = hs[e].energy_sum(tau_of(loop + 1), 1)
            _, Ws = hs[e].energy_sum(1.0, 2)
            balance.append((e, loop, (E1 - E0 + 0.5 * DT * (W0 + Ws)) / E0))
            energy[e].append((tau_of(loop), E0))
        batch.steps(ntskip - 1)
        dev_ms += batch.last_elapsed_ms()
        if single is not None:
            single.visc_steps(loop, ntskip)
            h = hs[check_event]
            check = {f: bool(np.array_equal(h.get(f), single.get(f))) for f in ('ev1', 'pi1', 'bulk1', 'udiff')}
            check['loops'] = loop + ntskip
            log('combined: event %d after %d steps equals its own handle: %s' % (check_event, loop + ntskip, check))
            single.close()
            single = None
        loop += ntskip
        changed = False
        for e in range(nev):
            if not active[e]:
                continue
            steps_done[e] += ntskip
            h = hs[e]
            h.hypersf_find(np.float32(tau_of(loop)))
            if h.max_ed(1) < efrz:
                active[e] = 0
                finish_tau[e] = tau_of(loop)
                changed = True
        if changed:
            batch.set_active(active)
        now = time.perf_counter()
        if now - t_last > 10.0 or not sum(active):
            t_last = now
            log('  tau %.2f fm (loop %d): %d events active, %d surface elements, %.1fs'
                % (tau_of(loop), loop, sum(active), sum(h.hypersf_count()[0] for h in hs), now - t_loop))
    wall = time.perf_counter() - t_loop
    updates = float(ncell) * sum(steps_done)
    value = updates / wall
    counts = [h.hypersf_count() for h in hs]
    nsf = sum(c[0] for c in counts)
    overflow = [e for e, c in enumerate(counts) if c[0] > c[1]]
    # emit the surface: copy each event's elements to the host (and write them when asked)
    t0 = time.perf_counter()
    dsum = 0.0
    for e, h in enumerate(hs):
        if counts[e][0] > counts[e][1]:
            continue
        sf, tx, pisf, bksf = h.hypersf_get()
        dsum += float(np.abs(sf[:, :4].astype(np.float64)).sum())
        if sf_out:
            from clvisc_amd import hypersf as hsf
            d = os.path.join(sf_out, 'event%d' % e)
            os.makedirs(d, exist_ok=True)
            hsf.write_hypersf(d, sf, tx, 0.137)
    t_emit = time.perf_counter() - t0
    res = np.array([b[2] for b in balance])
    log('combined: %d loops, %.1fs wall (%.1fs device steps) -> %.4g cell-updates/s; %d surface elements '
        '(copied in %.1fs); energy balance per step |res| max %.3g median %.3g'
        % (loop, wall, dev_ms / 1e3, value, nsf, t_emit, np.abs(res).max(), np.median(np.abs(res))))
    edmax_end = [h.max_ed(1) for h in hs]
    # the plain GPU form: the same events one after another, a window of steps each
    for e in range(nev):
        active[e] = 0
    batch.set_active(active)
    nwin = 10
    t0 = time.perf_counter()
    for e, h in enumerate(hs):
        h.visc_steps(loop, nwin, sync=False)
    for h in hs:
        h.sync()
    seq = float(ncell) * nev * nwin / (time.perf_counter() - t0)
    batch.close()
    return {'value': value, 'wall': wall, 'dev_ms': dev_ms, 'loops': loop, 'steps_done': steps_done,
            'finish_tau': finish_tau, 'seeds': seeds, 'surface_elements': nsf, 'surface_counts': counts,
            'surface_overflow_events': overflow, 'surface_abs_dsigma_sum': dsum, 'balance': res,
            'energy': energy, 'edmax_end': edmax_end, 'check': check,
            'step_value': updates / (dev_ms / 1e3), 'plain_sequential_value': seq}


def train_combined_seed(p, tab):
    """The seed network of the learned + batched combined run: 2000 x 4096 Adam steps on the
    states of the first six split steps of the first event (seed 20240601)."""
    from clvisc_amd import _lib, surrogate
    from clvisc_amd import ic as ic_mod
    h = _lib.Handle(p, tab, viscous=True)
    pos = ic_mod.woods_saxon_positions(400, 20240601)
    h.visc_initialize(TAU0)
    h.ideal_stage(1, np.float32(TAU0))
    h.set('ev1', ic_mod.hotspots(NX, NY, NZ, DX, DY, DZ, positions=pos, emax=50.0))
    t0 = time.perf_counter()
    states = surrogate.visc_run_states(h, 0, 6, every=3)
    h.close()
    mlp, hist = surrogate.train_seed(states, steps=2000, batch=4096, lr=3e-3, nmax=5_000_000)
    info = {'states': int(states[0].size), 'seconds': time.perf_counter() - t0,
            'loss_first': float(hist[:20].mean()), 'loss_last': float(hist[-20:].mean())}
    log('combined learned: seed trained on %d states of 6 split steps of event 0 in %.1fs, loss %.4g -> %.4g'
        % (info['states'], info['seconds'], info['loss_first'], info['loss_last']))
    return mlp, info


def run_combined(args, p, tab, name):
    """Combined workload (combined_run) as one JSON line; with --learned also the learned +
    batched form of the same run (the seed network trained first), reported against it."""
    nev = args.events
    learned = None
    if args.learned:
        mlp, info = train_combined_seed(p, tab)
        rl = combined_run(p, tab, nev, args.max_loops, args.sf_capacity, '', seed_mlp=mlp)
        mlp.close()
        learned = {'value': rl['value'], 'step_only_value': rl['step_value'], 'wall_s': rl['wall'],
                   'loops': rl['loops'], 'finish_tau': rl['finish_tau'], 'surface_elements': rl['surface_elements'],
                   'energy_balance_max_abs': float(np.abs(rl['balance']).max()),
                   'energy_balance_median_abs': float(np.median(np.abs(rl['balance']))),
                   'training': info}
    r = combined_run(p, tab, nev, args.max_loops, args.sf_capacity, args.sf_out)
    if learned is not None:
        learned['learned_over_plain'] = learned['value'] / r['value']
        learned['step_only_learned_over_plain'] = learned['step_only_value'] / r['step_value']
        log('combined learned + batched: %.4g cell-updates/s (steps only %.4g) against plain batched %.4g (%.4g): '
            '%.3fx' % (learned['value'], learned['step_only_value'], r['value'], r['step_value'],
                       learned['learned_over_plain']))
    value, wall, loop, res = r['value'], r['wall'], r['loops'], r['balance']
    step_value, seq, energy, seeds = r['step_value'], r['plain_sequential_value'], r['energy'], r['seeds']
    abytes = sum(ALGO_BYTES[True])
    out = {
        'metric': 'cell-updates/s, combined: %d batched 385x385x115 viscous+bulk events to freeze-out' % nev,
        'value': value, 'unit': 'cell-updates/s', 'n_gpus': 1,
        'steps': int(sum(r['steps_done'])), 'warmup': 0, 'ms_per_step': 1000.0 * wall / max(1, loop),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': value / REF_CELL_UPDATES,
        'dtype': 'f32 (f64 root for |T^ti| > 0.99 T^tt, f64 energy sums)', 'data': 'synthetic',
        'config': {'workload': '%d events 385x385x115, 400 Woods-Saxon hotspots each (seeds %d..%d), '
                               'eta/s 0.08, bulk zeta/s peak 0.05, crossover EOS, until max T < 0.137 GeV, '
                               'surface every 20 steps (skip 3,3,1)' % (nev, seeds[0], seeds[-1]),
                   'device': name, 'events': nev, 'cells_per_event': NX * NY * NZ, 'loops': loop,
                   'finish_tau': r['finish_tau'], 'surface_elements': r['surface_elements'],
                   'surface_overflow_events': r['surface_overflow_events'],
                   'surface_abs_dsigma_sum': r['surface_abs_dsigma_sum'],
                   'energy_balance_max_abs': float(np.abs(res).max()),
                   'energy_balance_median_abs': float(np.median(np.abs(res))),
                   'energy_first_last': [[energy[e][0][1], energy[e][-1][1]] for e in range(min(nev, 4))],
                   'step_only_value': step_value, 'plain_sequential_value': seq,
                   'batched_over_plain': step_value / seq, 'learned_batched': learned},
        'roofline': {'bound': 'hbm', 'kernel': 'batch_stage (both stages)',
                     'achieved': round(abytes * step_value / 1e9, 2), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(abytes * step_value / 1e9 / HBM_PEAK_GBS, 5), 'traffic': None,
                     'step_algorithmic_bytes_per_cell': abytes},
        'cpu_baseline': None,
    }
    print(json.dumps(out))


def run_surrogate(args, name):
    """Learned root-solve seed on the ideal Gubser workload (201 x 201 x 1, 200 steps): states
    from the plain run, device training (BF16 MFMA), then the seeded run; iterations per cell,
    field agreement with the plain run and with the analytic flow."""
    from clvisc_amd import ic, surrogate
    from clvisc_amd.eos import Eos
    nsteps = args.steps if args.steps != 100 else 200
    t0 = time.perf_counter()
    states, it_plain = surrogate.gubser_run_states(nsteps=nsteps, nx=201, every=1)
    t_plain = time.perf_counter() - t0
    log('plain run: %d steps in %.2fs (with state collection), %d states, mean iterations %.3f'
        % (nsteps, t_plain, states[0].size, it_plain.mean()))
    t0 = time.perf_counter()
    mlp, hist = surrogate.train_seed(states, steps=2000, batch=4096, lr=3e-3, nmax=5_000_000)
    t_train = time.perf_counter() - t0
    log('trained 2000 steps of 4096 in %.2fs: loss %.4g -> %.4g' % (t_train, hist[:20].mean(), hist[-20:].mean()))
    t0 = time.perf_counter()
    st_seed, it_seed, h_seed = surrogate.gubser_run_states(nsteps=nsteps, nx=201, every=1, seed_mlp=mlp,
                                                           return_handle=True)
    t_seed = time.perf_counter() - t0
    log('seeded run: %d steps in %.2fs, mean iterations %.3f' % (nsteps, t_seed, it_seed.mean()))
    eos = Eos('ideal_gas')
    ncell = 201 * 201
    e_p, e_s = states[3][-ncell:], st_seed[3][-ncell:]
    big = e_p > 1e-3 * e_p.max()
    field_rel = float((np.abs(e_s[big] - e_p[big]) / e_p[big]).max())
    tau = 1.0 + nsteps * 0.02
    ea, norm = ic.gubser_ideal(201, 201, 1, 0.08, 0.08, 0.15, tau, q=1.0)
    _, norm0 = ic.gubser_ideal(201, 201, 1, 0.08, 0.08, 0.15, 1.0, q=1.0, e0=float(eos.f_ed(0.3)))
    ea = ea[:, 0] * norm0
    x = (np.arange(201) - 100) * 0.08
    X, Y = np.meshgrid(x, x, indexing='ij')
    r = np.sqrt(X * X + Y * Y).ravel() < 4.0
    l1 = lambda e: float(np.abs(e[r] - ea[r]).sum() / ea[r].sum())
    log('field: max rel |e_seeded - e_plain| %.3g; L1 vs analytic (r < 4 fm) plain %.4g seeded %.4g'
        % (field_rel, l1(e_p), l1(e_s)))
    # plain vs seeded step time (device, 50 steps each, no state collection)
    h_seed.set_root_seed(None, track_iters=False)
    h_seed.ideal_steps(nsteps, 50, sync=True)
    ms_plain = h_seed.last_elapsed_ms() / 50
    h_seed.set_root_seed(mlp, track_iters=False)
    h_seed.ideal_steps(nsteps + 50, 50, sync=True)
    ms_seed = h_seed.last_elapsed_ms() / 50
    log('step time: plain %.4f ms, seeded %.4f ms' % (ms_plain, ms_seed))
    out = {
        'metric': 'root-solve iterations per cell, learned seed vs plain (ideal Gubser 201x201, 200 steps)',
        'value': float(it_seed.mean()), 'unit': 'iterations/cell', 'n_gpus': 1,
        'steps': nsteps, 'warmup': 0, 'ms_per_step': ms_seed,
        'higher_is_better': False, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32 (network bf16 MFMA, fp32 accumulate)', 'data': 'synthetic',
        'config': {'workload': 'ideal Gubser flow 201x201x1, q=1/fm, T0=0.3 GeV, dtau 0.02, 200 steps; MLP '
                               '4-64-64-2 tanh trained 2000 steps x 4096 on <=5e6 states of the plain run',
                   'device': name, 'plain_iterations': float(it_plain.mean()),
                   'seeded_max_iterations': int(it_seed.max()), 'plain_max_iterations': int(it_plain.max()),
                   'train_seconds': t_train, 'loss_first': float(hist[:20].mean()),
                   'loss_last': float(hist[-20:].mean()),
                   'max_rel_e_seeded_vs_plain': field_rel, 'l1_vs_analytic_plain': l1(e_p),
                   'l1_vs_analytic_seeded': l1(e_s), 'ms_per_step_plain': ms_plain},
    }
    print(json.dumps(out))
    h_seed.close()


def run_aux(args, p, tab, name):
    """The device kernels around the step, on the benchmark grid (385x385x115): initial
    conditions (optical Glauber, smearing of 20000 partons, entropy -> e), the vorticity
    tensors and omega^mu, the freeze-out surface with pi, Pi and the 32 vorticity tensors
    interpolated, one BulkInfo.get, and the Lambda polarization (thermal) of the surface
    at 1024 momenta.  Each call is timed by the wall clock with the device idle before
    and after it (best of 3 after one warm-up)."""
    from clvisc_amd import _lib
    from clvisc_amd import ic as ic_mod
    from clvisc_amd.config import read_config
    from clvisc_amd.eos import Eos
    from clvisc_amd.bulkinfo import BulkInfo
    from clvisc_amd.ideal import DeviceField
    from clvisc_amd.polarization import Polarization
    eos = Eos('crossover')
    ncell = NX * NY * NZ
    p.nxskip, p.nyskip, p.nzskip = 3, 3, 1          # as the combined run
    h = _lib.Handle(p, tab, viscous=True)

    def timed(fn, reps=3):
        fn()
        h.sync()
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            out = fn()
            h.sync()
            dt = time.perf_counter() - t0
            best = dt if best is None or dt < best else best
        return 1000.0 * best, out

    res = {}
    cfg = read_config()
    g = _lib.CvGlauber.from_cfg(cfg)
    res['glauber_ini_ms'], _ = timed(lambda: h.glauber_ini(g, 1))
    rng = np.random.default_rng(7)
    npart = 20000
    pt = rng.exponential(0.6, npart)
    phi = rng.uniform(0, 2 * np.pi, npart)
    y = rng.normal(0.0, 1.5, npart)
    mt = np.sqrt(pt * pt + 0.3 ** 2)
    x, yy = rng.normal(0, 2.5, npart), rng.normal(0, 2.5, npart)
    t, z = TAU0 * np.cosh(y), TAU0 * np.sinh(y)
    partons = np.column_stack([mt * np.cosh(y), pt * np.cos(phi), pt * np.sin(phi), mt * np.sinh(y), t, x, yy, z])
    res['smear_20000_partons_ms'], _ = timed(lambda: h.smear_partons(partons))
    res['sd_to_ed_ms'], _ = timed(lambda: h.sd_to_ed(eos.s, eos.ed, 1))
    # the evolution these act on: the bench's initial field, one ntskip of steps
    pos = ic_mod.woods_saxon_positions(400, 20240601)
    ev = ic_mod.hotspots(NX, NY, NZ, DX, DY, DZ, positions=pos, emax=50.0)
    h.visc_initialize(TAU0)
    h.ideal_stage(1, np.float32(TAU0))
    h.set('ev1', ev)
    del ev
    h.vorticity_enable()
    h.hypersf_reserve(args.sf_capacity)
    h.hypersf_snapshot(np.float32(TAU0))
    nsteps = 20
    h.visc_steps(0, nsteps)
    tau = np.float32(TAU0 + nsteps * DT)
    res['vorticity_ms'], _ = timed(lambda: h.vorticity(tau))
    t0 = time.perf_counter()
    h.hypersf_find(tau)
    h.sync()
    res['hypersf_find_ms'] = 1000.0 * (time.perf_counter() - t0)
    nsf, _ = h.hypersf_count()
    res['vorticity_mu_with_copy_ms'], _ = timed(lambda: h.vorticity_mu(p.efrz), reps=1)
    c = read_config()
    c.NX, c.NY, c.NZ, c.DX, c.DY, c.DZ = NX, NY, NZ, DX, DY, DZ
    bi = BulkInfo(c, eos, device_reductions=True)
    d_ev1 = DeviceField(h, 'ev1')
    res['bulkinfo_get_ms'], _ = timed(lambda: bi.get(float(tau), d_ev1, 1.0))
    sf, _, _, _ = h.hypersf_get()
    om = h.hypersf_get_vorticity()
    h.close()
    nmom = 1024
    pz = Polarization.from_arrays(sf, omega=np.ascontiguousarray(om[:, 0:6]), T=0.137, mass=1.116)
    ptm = np.repeat(np.linspace(0.2, 3.0, 32), 32)
    phm = np.tile(np.linspace(0, 2 * np.pi, 32, endpoint=False), 32)
    mtm = np.sqrt(ptm ** 2 + 1.116 ** 2)
    pz.set_momentum(np.column_stack([mtm, np.zeros(nmom), ptm * np.cos(phm), ptm * np.sin(phm)]))
    pz.get_themal()
    t0 = time.perf_counter()
    pz.get_themal()
    res['polarization_thermal_ms'] = 1000.0 * (time.perf_counter() - t0)
    res['surface_elements'] = int(nsf)
    res['momenta'] = nmom
    for k_, v_ in res.items():
        log('  %-28s %s' % (k_, ('%.3f' % v_) if isinstance(v_, float) else v_))
    out = {
        'metric': 'ms per call of the device kernels around the viscous step (385x385x115)',
        'value': res['vorticity_ms'], 'unit': 'ms', 'n_gpus': 1, 'steps': nsteps, 'warmup': 1,
        'ms_per_step': None, 'higher_is_better': False, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': 'initial conditions, vorticity, freeze-out surface, BulkInfo and Lambda '
                               'polarization on the 385x385x115 grid after %d viscous + bulk steps' % nsteps,
                   'cells': ncell, 'device': name, 'timings': res,
                   'cells_per_s': {k_: ncell / (v_ * 1e-3) for k_, v_ in res.items()
                                   if k_ in ('glauber_ini_ms', 'sd_to_ed_ms', 'vorticity_ms')},
                   'element_momenta_per_s': nsf * nmom / (res['polarization_thermal_ms'] * 1e-3)},
    }
    print(json.dumps(out))


def run_counter_pass(args, h):
    """--warmup + 1 steps of the single workload, then one device-to-device copy of a
    pi-sized field (NX NY NZ x 10 floats = 681.8 MB read and written, cv_copy_field): the
    kernel of known size that scales FETCH_SIZE / WRITE_SIZE to bytes."""
    h.visc_steps(0, args.warmup + 1)
    h.sync()
    h.copy('pi2', 'pi1')
    h.sync()
    log('counter pass: %d steps, then a %.1f MB field copy' % (args.warmup + 1, 40.0 * NX * NY * NZ / 1e6))
    h.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=100)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--no-bulk', action='store_true', help='shear viscosity only')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-learned', action='store_true',
                    help='single: skip the learned (network-seeded root solve) form of the run')
    ap.add_argument('--cpu-threads', type=int, default=0,
                    help='threads of the CPU baseline (default: OMP_NUM_THREADS, else the usable CPUs)')
    ap.add_argument('--no-c1', action='store_true', help='single: skip the config-1 (ideal Gubser 201x201) line')
    ap.add_argument('--events', type=int, default=16, help='combined: number of 385x385x115 events')
    ap.add_argument('--max-loops', type=int, default=3000, help='combined: step limit')
    ap.add_argument('--sf-capacity', type=int, default=6_000_000, help='combined: surface elements per event')
    ap.add_argument('--sf-out', default='', help='combined: write hypersf.dat / sf_txyz.dat per event here')
    ap.add_argument('--workload', choices=['single', 'ensemble', 'surrogate', 'combined', 'aux'], default='single',
                    help='single: one 385x385x115 event (the headline); ensemble: 64 uneven events batched; '
                         'surrogate: learned root-solve seed on the ideal Gubser workload')
    ap.add_argument('--learned', action='store_true',
                    help='ensemble, combined: also the learned + batched form (seeded root solve) of the run')
    ap.add_argument('--autotune', action='store_true', help='ensemble: tune tiles and chunks per grid shape')
    ap.add_argument('--autotune-cache', default='', help='ensemble: JSON file of tuned shapes')
    ap.add_argument('--counter-pass', action='store_true',
                    help='single: only --warmup + 1 steps, then one 682 MB device copy (the byte '
                         'counters\' calibration kernel); for rocprofv3 --pmc passes')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU only: the hydro step of one event runs on one MI355X '
              '(multi-GPU slab decomposition is described in DESIGN.md, not built)', file=sys.stderr)
        sys.exit(2)

    from clvisc_amd import _lib
    bulk = not args.no_bulk
    if args.workload == 'surrogate':
        name, cus = _lib.device_info(0)
        log('device: %s, %d CUs' % (name, cus))
        run_surrogate(args, name)
        return
    if args.workload == 'combined':
        p, tab, _ = make_workload(bulk, ic=False)
        name, cus = _lib.device_info(0)
        log('device: %s, %d CUs' % (name, cus))
        run_combined(args, p, tab, name)
        return
    if args.workload == 'aux':
        p, tab, _ = make_workload(bulk, ic=False)
        name, cus = _lib.device_info(0)
        log('device: %s, %d CUs' % (name, cus))
        run_aux(args, p, tab, name)
        return
    if args.workload == 'ensemble':
        p, tab, _ = make_workload(bulk, ic=False)
        name, cus = _lib.device_info(0)
        log('device: %s, %d CUs' % (name, cus))
        run_ensemble(args, p, tab, name)
        return
    p, tab, ev = make_workload(bulk)
    name, cus = _lib.device_info(0)
    log('device: %s, %d CUs' % (name, cus))
    h = _lib.Handle(p, tab, viscous=True)
    h.visc_initialize(TAU0)
    h.ideal_stage(1, np.float32(TAU0))
    h.set('ev1', ev)
    if args.counter_pass:
        run_counter_pass(args, h)
        return
    loop = 0
    h.visc_steps(loop, args.warmup)
    loop += args.warmup
    # per-launch device time of one step (HIP events on the handle's stream)
    prof = h.profile_visc_step(loop)
    loop += 1
    h.sync()
    t0 = time.perf_counter()
    h.visc_steps(loop, args.steps, sync=False)
    h.sync()
    wall = time.perf_counter() - t0
    dev_ms = h.last_elapsed_ms()
    loop += args.steps
    ncell = NX * NY * NZ
    ms_per_step = 1000.0 * wall / args.steps
    value = ncell * args.steps / wall
    edmax = h.max_ed(1)
    kind, chunk = h.stage_kernel(), h.march_chunk()
    # the hotspot field only decays (peak 50 GeV/fm^3 at tau0): a larger or non-finite max
    # means the evolution went unstable and the timing is not of the physical workload
    ok = bool(np.isfinite(edmax) and 0.0 < edmax <= 50.0)
    if not ok:
        log('WARNING: max e after the run is %r: the evolution is not physical' % edmax)
    log('steps %d wall %.3fs device %.3f ms -> %.3f ms/step, %.4g cell-updates/s, edmax %.4g'
        % (args.steps, wall, dev_ms, ms_per_step, value, edmax))
    abytes = ALGO_BYTES[bulk]
    per = [(abytes[i] * ncell / (prof[i] * 1e-3) / 1e9) for i in range(len(prof))]
    for i, nm in enumerate(LAUNCH_NAMES):
        log('  %-10s %8.3f ms  %6.1f GB/s algorithmic (%d B/cell)' % (nm, prof[i], per[i], abytes[i]))
    dom = int(np.argmax(prof))
    roof = {'bound': 'hbm', 'kernel': LAUNCH_NAMES[dom], 'achieved': round(per[dom], 2),
            'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(per[dom] / HBM_PEAK_GBS, 5),
            'traffic': profiled_traffic(LAUNCH_NAMES[dom]) if bulk else None,
            'step_algorithmic_bytes_per_cell': sum(abytes),
            'step_frac': round(sum(abytes) * value / 1e9 / HBM_PEAK_GBS, 5)}
    # what bounds the kernel: its VALU instructions per dispatch (SQ_INSTS_VALU of the
    # committed profile) issued in this run's dispatch time, against the VALU issue rate
    valu = profiled_value('valu_insts', LAUNCH_NAMES[dom]) if bulk else None
    if valu:
        roof['valu_issue_frac'] = round(valu / VALU_ISSUE_PER_S / (prof[dom] * 1e-3), 4)
    learned = None
    if not args.no_learned:
        learned = run_learned(args, p, tab, ev, h, loop)
        learned['learned_over_plain'] = learned['value'] / value
        learned['split_learned_over_plain'] = learned['split_value'] / value
    cpu = None
    nth = args.cpu_threads or cpu_threads_default()
    if not args.no_cpu_baseline:
        t0 = time.time()
        cpu = cpu_baseline(p, tab, ev, nth)
        log('cpu baseline: %.4g cell-updates/s on %d threads (%.1f ms/step), %.4g on 1 thread (slab) (%.1fs)'
            % (cpu['value'], nth, cpu['ms_per_step'], cpu['one_thread']['value'], time.time() - t0))
        log('host: %s' % json.dumps(cpu['host']))
    c1 = None
    if not args.no_c1:
        t0 = time.time()
        c1 = c1_gubser(nth, with_cpu=not args.no_cpu_baseline)
        log('C1 ideal Gubser 201x201: GPU %.4f ms/step (%.4g cell-updates/s), L1(e) vs analytic %.4f%s (%.1fs)'
            % (c1['gpu_ms_per_step'], c1['gpu_value'], c1['l1_e_vs_analytic'],
               '' if 'cpu' not in c1 else '; CPU %.3f ms/step on %d threads, %.3f on 1 (bit-equal to the GPU: %s)'
               % (c1['cpu']['ms_per_step'], nth, c1['cpu_one_thread']['ms_per_step'], c1['cpu']['bit_equal_gpu']),
               time.time() - t0))
    out = {
        'metric': 'cell-updates/s, viscous KT+IS 385x385x115; learned vs plain at matched error',
        'value': value, 'unit': 'cell-updates/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': ms_per_step,
        'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': value / REF_CELL_UPDATES,
        'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': 'viscous KT + Israel-Stewart shear (eta/s=0.08)%s, 385x385x115, '
                               'dtau=0.02 fm, 400 Woods-Saxon hotspots, crossover EOS (octave table)'
                               % (' + bulk (zeta/s peak 0.05 at 0.155 GeV)' if bulk else ''),
                   'grid': [NX, NY, NZ], 'device': name, 'physical': ok,
                   'edmax_end': float(edmax),
                   'stage_kernel': {'kind': kind[0], 'tile': [kind[1], kind[2]],
                                    'march_chunk': chunk[0], 'chunks': chunk[1]}},
        'roofline': roof,
        'learned': learned,
        'cpu_baseline': cpu,
        'c1_gubser': c1,
    }
    print(json.dumps(out))
    h.close()


if __name__ == '__main__':
    main()
