/*
 * clvisc_amd.h -- C-ABI of the MI355X (gfx950) KT + Israel-Stewart hydro step.
 *
 * This is the drop-in boundary for the reference's hydro path.  In the reference,
 * Python driver classes (pyvisc/ideal.py CLIdeal, pyvisc/visc.py CLVisc) enqueue
 * OpenCL kernels through pyopencl; here the same drivers (clvisc_amd.CLIdeal /
 * clvisc_amd.CLVisc) call these entry points through ctypes.  Each entry point
 * below names the reference code it replaces.
 *
 * Conventions
 *   - Every call returns 0 on success and a negative CV_ERR_* code on failure;
 *     cv_last_error() returns a message for the calling thread.
 *   - Calls are synchronous at return unless named *_async.
 *   - Host arrays are caller-owned, contiguous, in the reference's AoS layouts:
 *       ev    float[N][4]   (e, vx, vy, veta)  with veta = tau * deta/dtau
 *       pi    float[N][10]  pi^{mu nu} in the order 00,01,02,03,11,12,13,22,23,33,
 *                           eta components tilde-scaled (tau pi^{mu eta}, tau^2 pi^{eta eta})
 *       N = NX*NY*NZ, cell I = (i*NY + j)*NZ + k.
 *     The device keeps fields as structure-of-arrays planes and converts in get/set.
 *   - Arithmetic is FP32 and follows one contract: a/b is evaluated as a * RN(1/b),
 *     sqrt is correctly rounded, no FMA contraction.  With the octave EOS table the
 *     results are reproducible bit for bit by an IEEE CPU evaluating the same
 *     operations.
 */
#ifndef CLVISC_AMD_H
#define CLVISC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CV_OK              0
#define CV_ERR_ARG        -1
#define CV_ERR_HIP        -2
#define CV_ERR_STATE      -3
#define CV_ERR_CAPACITY   -4
#define CV_ERR_UNSUPPORTED -5

/* flags of cv_params.flags (compile-time -D switches of the reference) */
#define CV_FLAG_RIEMANN_TEST     0x1   /* -D RIEMANN_TEST (pyvisc/ideal.py:164-165) */
#define CV_FLAG_GUBSER_VISC_TEST 0x2   /* -D GUBSER_VISC_TEST (pyvisc/visc.py:111-112) */
#define CV_FLAG_BULK             0x4   /* bulk viscous pressure Pi (no reference counterpart) */
#define CV_FLAG_FP64_ROOT        0x8   /* FP64 root solve where M > fp64_root_vmin * T00 */
#define CV_FLAG_SF_FP64          0x10  /* freeze-out surface geometry in FP64 (exact where FP32
                                          loses the tie-breaking shift of coplanar intersections) */
#define CV_FLAG_PIMN_OMEGA       0x20  /* -D PIMUNU_OMEGA_COUPLING (pyvisc/visc.py:117-118) */
#define CV_FLAG_OMEGA_OMEGA      0x40  /* -D OMEGA_OMEGA_COUPLING (pyvisc/visc.py:120-121) */

/* Hydro parameters: the reference's -D compile options (pyvisc/ideal.py:146-171,
 * pyvisc/visc.py:108-121) as run-time values. */
typedef struct cv_params {
    int32_t nx, ny, nz;
    int32_t flags;
    double  tau0, dt;                 /* TAU0, DT as the config holds them */
    float   dx, dy, dz;
    float   etaos_xmin, etaos_ymin, etaos_left_slop, etaos_right_slop;
    float   lam1;
    float   zetaos_max, zetaos_tpeak, zetaos_width;   /* CV_FLAG_BULK */
    float   fp64_root_vmin;                           /* CV_FLAG_FP64_ROOT */
    float   efrz;                                     /* freeze-out energy density */
    int32_t nxskip, nyskip, nzskip;
    int32_t reserved;
} cv_params;

/* EOS table (replaces Eos.create_table's image2d, pyvisc/eos/eos.py:184-203).
 * kind 0: the reference's uniform row/col table (eos_table.h:39-100).
 * kind 1: octave table, nodes 2^k (1 + j/2^log2j), k in [kmin, kmax), plus 2^kmax. */
typedef struct cv_eos {
    int32_t kind;
    int32_t nrow, ncol, num_ed;       /* kind 0 */
    float   ed_start, ed_step;        /* kind 0 */
    int32_t kmin, kmax, log2j;        /* kind 1 */
    int32_t n_entries;                /* number of RGBA entries in rgba */
    const float *rgba;                /* host (cs2, P, T, s) x n_entries, copied at create */
} cv_eos;

typedef struct cv_handle cv_handle;

/* Field ids for get/set/copy (the reference's device buffers) */
enum cv_field {
    CV_EV0 = 0, CV_EV1 = 1, CV_EV2 = 2,        /* ideal.d_ev[0..2]   (ideal.py:102-104) */
    CV_PI0 = 3, CV_PI1 = 4, CV_PI2 = 5,        /* visc.d_pi[0..2]    (visc.py:39-41)    */
    CV_BULK0 = 6, CV_BULK1 = 7, CV_BULK2 = 8,  /* bulk Pi (extension), float[N]          */
    CV_SRC = 9,                                /* ideal.d_Src        (ideal.py:105)     */
    CV_IS_SRC = 10,                            /* visc.d_IS_src      (visc.py:43)       */
    CV_UDIFF = 11,                             /* visc.d_udiff       (visc.py:83)       */
    CV_GOODCELL = 12,                          /* visc.d_goodcell    (visc.py:87-88)    */
    CV_ROOT_CONS = 13,                         /* split root solve: its conserved input
                                                  (T^tt, T^tx, T^ty, T^t eta), no counterpart */
    CV_ROOT_SEED = 14,                         /* split root solve: the learned seed |v|    */
    CV_OMEGA0 = 15, CV_OMEGA1 = 16,            /* visc.d_omega[0..1] + d_omega_shear1/2 +
                                                  d_omega_accT (visc.py:52-71), 32 per cell */
    CV_NFIELDS = 17
};

/* Optical Glauber model (pyvisc/kernel/kernel_glauber.cl, pyvisc/glauber.py): the
 * [glauber] section of hydro.info.  bjorken = 1 is the BJORKEN_SCALING define (H(eta) = 1). */
typedef struct cv_glauber {
    float num_nucleons;     /* NumOfNucleons */
    float sqrts;            /* SQRTS (GeV) */
    float ro0;              /* NucleonDensity (fm^-3) */
    float ra;               /* Ra (fm) */
    float eta;              /* Eta: Woods-Saxon diffuseness (fm) */
    float si0;              /* Si0: inelastic cross section (fm^2) */
    float b;                /* ImpactParameter (fm) */
    float edmax;            /* Edmax (GeV/fm^3) at b = 0 */
    float hwn;              /* Hwn: wounded-nucleon fraction */
    float eta_flat;         /* Eta_flat */
    float eta_gw;           /* Eta_gw */
    int32_t bjorken;
} cv_glauber;

/* Last error message of the calling thread (never NULL). */
const char *cv_last_error(void);

/* Library / device info: fills name (>= 64 bytes) and compute units. */
int cv_device_info(int device, char *name, int name_len, int32_t *cus);

/* No counterpart in the reference.  Device check of the two short forms the stage kernels use
 * where the operand is known to lie in their range: the reciprocal (hardware reciprocal + one
 * FMA Newton step) against the correctly rounded 1/b for every float with
 * 2^-126 <= |b| < 2^126, and the square root (hardware reciprocal square root + one FMA
 * Newton step) against the correctly rounded sqrtf for every float of [2^-24, 1], bit for
 * bit.  out[0] / out[1]: reciprocal mismatches / floats compared; out[2] / out[3]: square
 * root mismatches / floats compared (zero mismatches are required for the stage kernels'
 * results to equal the oracle's). */
int cv_selftest_fastmath(int device, uint64_t out[4]);

/* Construction = CLIdeal.__init__ / CLVisc.__init__ buffer setup (ideal.py:30-119,
 * visc.py:24-104) on HIP device `device`.  Allocates all fields zeroed; viscous != 0
 * also allocates pi / IS_src / udiff / goodcell.  Does not run any kernel. */
int cv_create(const cv_params *p, const cv_eos *eos, int device, int viscous, cv_handle **out);
int cv_destroy(cv_handle *h);

/* Host <-> device copies of a field in the reference's AoS layout
 * (cl.enqueue_copy on d_ev[k] / d_pi[k], e.g. load_ini ideal.py:138-144). */
int cv_set_field(cv_handle *h, int field, const float *host, int64_t nfloats);
int cv_get_field(cv_handle *h, int field, float *host, int64_t nfloats);
int cv_copy_field(cv_handle *h, int dst, int src);            /* device-to-device */

/* Rows of field at the given cell indices, AoS [ncells][ncomp] into host: the slice
 * kernels of BulkInfo / PimnInfo (pyvisc/bulkinfo.py:66-126 get_ed / get_pimn,
 * pyvisc/pimninfo.py:41-67) without copying the whole field. */
int cv_gather(cv_handle *h, int field, const int64_t *cells, int64_t ncells, float *host);

/* visc_initialize kernel (kernel_IS.cl:10-36) as CLVisc.IS_initialize (visc.py:143-149). */
int cv_visc_initialize(cv_handle *h, float tau);

/* CLIdeal.stepUpdate(step) at proper time tau (ideal.py:211-255): kt_src_christoffel,
 * kt_src_along{x,y,z}, update_ev of kernel_ideal.cl, fused into one launch. */
int cv_ideal_stage(cv_handle *h, int step, float tau);

/* CLVisc.IS_stepUpdate(step) (visc.py:287-314): visc_src_christoffel,
 * visc_src_along{x,y,z}, update_pimn of kernel_IS.cl, fused into one launch. */
int cv_is_stage(cv_handle *h, int step, float tau);

/* CLVisc.visc_stepUpdate(step) (visc.py:232-283): kt_src_christoffel, kt_src_along*,
 * update_ev of kernel_visc.cl, fused into one launch. */
int cv_visc_stage(cv_handle *h, int step, float tau);

/* CLVisc.update_udiff(d_ev_a, d_ev_b) (visc.py:352-357, get_udiff kernel_IS.cl:714-724). */
int cv_update_udiff(cv_handle *h, int ev_a, int ev_b);

/* nsteps full RK2 steps of the evolve loops starting at loop index loop0
 * (ideal: ideal.py:341-344; viscous: visc.py:527-544, including the d_pi[0]/d_ev[0]
 * copies and update_udiff).  tau after the call is TAU0 + (loop0+nsteps)*DT rounded
 * to float.  Work is queued on the handle's stream; the call returns once queued. */
int cv_ideal_steps_async(cv_handle *h, int64_t loop0, int64_t nsteps);
int cv_visc_steps_async(cv_handle *h, int64_t loop0, int64_t nsteps);

/* max of e over d_ev[which] (reduction_stage1 + host max, kernel_reduction.cl:7-39,
 * ideal.py:257-262). */
int cv_max_ed(cv_handle *h, int which, float *out);

/* pizz_o_ep of CLVisc.check_pizz (kernel_checkpi.cl:5-27, visc.py:200-208): out[K] (NZ floats,
 * host memory) = mean over the cells of eta slice K with e + P > 0.133 of pi^{eta eta}/(e + P),
 * from d_ev[which] and d_pi[which], summed in the reference's cell order.  Viscous handles
 * only (CV_ERR_UNSUPPORTED otherwise).  Synchronous. */
int cv_pizz_o_ep(cv_handle *h, int which, float *out);

/* FP64 energy balance of slot `which` at proper time tau (no counterpart in the reference):
 * out[0] = sum over cells of tau (T^tautau + pi^tautau) dx dy deta (the conserved Q^tau of
 * kernel_visc.cl:322-393 summed), out[1] = sum of T~^etaeta = tau^2 T^etaeta (+ pi~^etaeta)
 * dx dy deta, the only non-flux source of Q^tau (kt_src_christoffel, kernel_visc.cl:11-40):
 * dE/dtau = -W plus the flux through the grid boundary.  Summed in double from the FP32
 * fields in a fixed order (reproducible). */
int cv_energy_sum(cv_handle *h, int which, double tau, double *out);

/* BulkInfo observables of slot `which` per eta slice, replacing the whole-field copy of
 * BulkInfo.ecc_vs_rapidity (pyvisc/bulkinfo.py:247-260), the plane sums of eccp /
 * mean_vr / total_entropy (bulkinfo.py:203-245) and the total_energy_and_entropy and
 * eccp_vs_rapidity kernels (kernel_bulkinfo.cl:18-113, bulkinfo_h5.py:264-297):
 * out[k * 10 + q] = sum over the (x, y) plane at eta index k of q = 0: Txx - Tyy,
 * 1: Txx + Tyy, 2: T^tx, 3: u0 s, 4: u0 e |v_T|, 5: u0 e (e floored at 1e-10, v^2 > 1 ->
 * 0.999999 as bulkinfo.py does), and, with kernel_bulkinfo.cl's unfloored e and gamma
 * (1 - v^2 floored at 1e-7): 6: (e + P) u0^2, 7: (e + P) u0^2 v^eta, 8: P, 9: u0 s; the
 * slice's lab-frame energy density sum is cosh(eta_k) (q6 - q8) + sinh(eta_k) q7.  Double
 * sums, fixed order. */
int cv_slice_sums(cv_handle *h, int which, double *out);
/* No counterpart in the reference.  cv_slice_sums for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the same kernels and sums, with the partial sums in a static device array of
 * the library (at most 128 eta slices; 10.5 MB per GPU, held until the process ends) instead
 * of a buffer allocated on the first call.  d_out: caller-owned device memory of nz * 10
 * doubles, written on the handle's stream; the call returns without waiting (cv_sync).
 * Grids of more than 128 eta slices fail with CV_ERR_CAPACITY.  Calls that overlap on one GPU
 * share the array and race. */
int cv_slice_sums_noalloc(cv_handle *h, int which, double *d_out);

/* ---- initial conditions on the device ------------------------------------------------
 * glauber_ini (kernel_glauber.cl:82-102, Glauber.__init__ pyvisc/glauber.py:63-70): slot
 * `which` <- (Edmax e_T(x, y) / e_T(0, 0) H(eta_s), 0, 0, 0). */
int cv_glauber_ini(cv_handle *h, const cv_glauber *g, int which);
/* num_of_binary_collisions (kernel_glauber.cl:104-121, Glauber.save_nbinary
 * pyvisc/glauber.py:95-105): out[i * ny + j] = Si0 T(x - b/2, y) T(x + b/2, y) on the grid
 * x = xlow + i dx, y = ylow + j dy (host array of nx * ny floats). */
int cv_glauber_nbinary(cv_handle *h, const cv_glauber *g, float xlow, float ylow, float dx, float dy, int nx,
                       int ny, float *out);
/* Gaussian smearing of partons (smearing kernel, kernel_gaussian_smearing_new.cl:11-120;
 * SmearingP4X4 pyvisc/smearing.py:63-90): p4x4 = npartons rows of (E, px, py, pz, t, x, y,
 * z) (the reference's real8); slot `which` <- (e, 0, 0, 0) with e from the smeared
 * T^{tau mu} by the fixed-point root solve of helper.h:112-125. */
int cv_smear_partons(cv_handle *h, const float *p4x4, int64_t npartons, float sigr, float sigz, float kfactor,
                     int which);
/* change_longitudinal_profile (kernel_gaussian_smearing_new.cl:124-138): every cell of an
 * (x, y) column <- (profile[k] e_mid, v_mid) of its mid-rapidity cell; profile: NZ floats. */
int cv_longitudinal_profile(cv_handle *h, int which, const float *profile);
/* Entropy density -> energy density in place, component e of slot `which` (from_sd_to_ed,
 * pyvisc/ebe_trento.py:26-33): piecewise linear through the m EOS nodes (s_nodes strictly
 * increasing, e_nodes), linear extrapolation at both ends, in double. */
int cv_sd_to_ed(cv_handle *h, int which, const double *s_nodes, const double *e_nodes, int64_t m);

/* ---- vorticity ------------------------------------------------------------------------
 * Allocate d_omega[0..1] (32 components per cell: omega^{tx,ty,tz,xy,xz,yz}, omega_shear1
 * (16), omega_shear2 (4), omega_accT (6); visc.py:52-71) and, with the surface, 32 floats
 * per surface element.  Viscous handles only. */
int cv_vorticity_enable(cv_handle *h);
/* get_vorticity (visc.py:316-335): kernel_vorticity.cl omega / omega_shear / omega_accT from
 * d_ev[0] (previous) and d_ev[1] at tau into d_omega[1].  After each cv_hypersf_find the
 * tensors are interpolated to the elements and d_omega[0] <- d_omega[1] (visc.py:383-395). */
int cv_vorticity(cv_handle *h, float tau);
/* omega_mu (kernel_vorticity.cl:103-140, visc.py:336-350): host AoS rows (T, omega^x,
 * omega^y, omega^eta) from d_ev[1] and d_omega[1], zero where e < 0.1 efrz. */
int cv_vorticity_mu(cv_handle *h, float efrz, float *host, int64_t nfloats);
/* The 32 vorticity components of each stored surface element (omegamu_sf.dat,
 * omegamu_shear1_sf.dat, omegamu_shear2_sf.dat, omegamu_accT_sf.dat of visc.py:430-458). */
int cv_hypersf_get_vorticity(cv_handle *h, float *om, int64_t nelem);

/* Lambda polarization from the freeze-out surface (pyvisc/kernel/kernel_polarization.cl,
 * Polarization.get_themal / get_shear / get_accT of pyvisc/polarization/polarization.py):
 * sf = nsf rows of 8 (hypersf.dat: dS_mu, v, eta_s); kind 0: om1 = omegamu_sf rows (6);
 * kind 1: om1 = omegamu_shear1_sf rows (16), om2 = omegamu_shear2_sf rows (4); kind 2:
 * om1 = omegamu_accT_sf rows (6).  mom = nmom rows (m_T, Y, p_x, p_y).  Outputs per
 * momentum: pol_lab (4, = sum / (4 mass)), density, pol_lrf (4, rest frame of the Lambda).
 * No handle: device index only. */
int cv_polarization(int device, int kind, const float *sf, const float *om1, const float *om2, int64_t nsf,
                    const float *mom, int64_t nmom, float tfrz, float mass, float *pol_lab, float *density,
                    float *pol_lrf);

/* Profiling: runs one viscous step (as cv_visc_steps_async with nsteps = 1) with HIP
 * events around each of its launches, waits, and writes the launch durations in
 * milliseconds to ms[0..nms-1] in launch order; returns the number of launches in
 * *nlaunch.  Same results as the unprofiled step. */
int cv_profile_visc_step(cv_handle *h, int64_t loop0, float *ms, int nms, int32_t *nlaunch);

/* Kernel of the step path (cv_visc_steps_async): kind 0 = one thread per cell, every
 * stencil quantity re-evaluated per cell; kind 1 = tiled kernel, one workgroup per
 * ty x tz tile of an x-plane, per-cell quantities and y/z faces evaluated once in LDS;
 * kind 2 (default) = the tiled kernel marching along x over chunks of planes, every x face
 * and slope evaluated once (ty, tz <= 0 picks the default tile for the grid).  The ideal
 * step (cv_ideal_steps_async, cv_ideal_stage) runs on x-marching tiles of the same shape
 * with kind 2 and per cell otherwise (also with the test flags, CV_FLAG_FP64_ROOT, a root
 * seed or iteration counts).  All give the same FP32 results.  No reference counterpart
 * (the reference's launch geometry is BSZ, pyvisc/hydro.info). */
int cv_set_stage_kernel(cv_handle *h, int kind, int ty, int tz);
/* Jet energy-loss source (jet_eloss_src of pyvisc/kernel/kernel_jet_eloss.cl:8-69, passed as
 * the dict {'switch_on', 'direction', 'start_pos_index'} to CLIdeal.stepUpdate / CLVisc.evolve,
 * ideal.py:211-246, visc.py:232-271): on is a mask of the RK stages that add it (1: stage 1,
 * 2: stage 2, 3: both; 0: none).  The reference's evolve passes it to stage 2 only
 * (ideal.py:344, visc.py:543): on = 2 for the steps calls.  Each such stage adds, after its KT
 * sources, a Gaussian (0.3 fm) deposit of dE/dx = 14 GeV/fm (T/0.3 GeV)^3 around a point that
 * leaves cell start_pos_index (i NY NZ + j NZ + k) at tau0 along (cos direction, sin direction, 0)
 * at the speed of light.  With a jet set, stages run in the per-cell kernels.  CV_ERR_ARG for a
 * start outside the grid or a mask above 3, CV_ERR_UNSUPPORTED with cv_set_root_split. */
int cv_set_jet_source(cv_handle *h, int on, float direction, int64_t start_pos_index);
int cv_get_stage_kernel(cv_handle *h, int32_t *kind, int32_t *ty, int32_t *tz);
/* x planes per march chunk of kind 2 (0: automatic, about 32 work items per CU). */
int cv_set_march_chunk(cv_handle *h, int xc);
/* xc (the longest chunk) and the number of chunks along x (automatic chunks cut per XCD
 * slab: nxcd slabs of equally many chunks, see cv_set_xcd_map). */
int cv_get_march_chunk(cv_handle *h, int32_t *xc, int32_t *nchunks);
/* Work mapping of the tiled stage kernels (no counterpart in the reference): nxcd > 0 gives
 * each of the nxcd XCDs, to which the hardware hands workgroups round-robin, a contiguous
 * range of work items (neighbouring tiles share an L2); 0 keeps workgroup b on item b.
 * With automatic march chunks of at least twice their tail, nxcd > 1 and nx >= 2 nxcd, the
 * range of an XCD is one slab of x planes, cut into chunks with a tail of 1 to 4 planes
 * last, so that the items an XCD runs last are short ones.  Default 8 (MI355X has 8 XCDs).
 * Results are the same either way. */
int cv_set_xcd_map(cv_handle *h, int nxcd);

/* Wait for all queued work. */
int cv_sync(cv_handle *h);

/* Device timing of the last cv_*_steps_async call bracketed with HIP events:
 * elapsed milliseconds between its first and last launch. */
int cv_last_elapsed_ms(cv_handle *h, float *ms);

/* ---- freeze-out hypersurface (pyvisc/kernel/kernel_hypersf.cl:405-660) ----------
 * Cubes of the grid thinned by (nxskip, nyskip, nzskip) x (tau_old, tau_new) where
 * e crosses cv_params.efrz.  Elements are appended in an order that is not defined
 * (atomic append, as the reference's atomic_inc): per element
 *   sf    float[8]  dSigma_0..3, vx, vy, veta, eta_s   (hypersf.dat rows)
 *   txyz  float[4]  t, x, y, z                          (sf_txyz.dat rows)
 *   pisf  float[10] pi^{mu nu} at the centroid          (viscous; pimnsf.dat rows)
 *   bksf  float     bulk pressure at the centroid       (CV_FLAG_BULK only)
 * Elements beyond the capacity are counted but not stored; cv_hypersf_get then fails
 * with CV_ERR_CAPACITY (the reference writes past its 1.5M-element buffers). */

/* (Re)allocate the element buffers for capacity elements (default 1,500,000 at the
 * first snapshot, the size of the reference's d_hypersf, pyvisc/ideal.py:113-117).
 * Keeps the count. */
int cv_hypersf_reserve(cv_handle *h, int64_t capacity);

/* d_ev_old <- d_ev[1] (viscous: d_pi_old <- d_pi[1]; bulk: its previous field too),
 * tau_old <- tau: the n == 0 branch of get_hypersf (pyvisc/ideal.py:275-278,
 * pyvisc/visc.py:367-371). */
int cv_hypersf_snapshot(cv_handle *h, float tau);

/* get_hypersf / visc_hypersf between the snapshot and d_ev[1] at tau_new, then the
 * snapshot update (pyvisc/ideal.py:279-291, pyvisc/visc.py:372-397). */
int cv_hypersf_find(cv_handle *h, float tau_new);

/* Elements found since creation or the last reset (may exceed *capacity), the
 * reference's d_num_of_sf read back in save (pyvisc/ideal.py:295-300). */
int cv_hypersf_count(cv_handle *h, int64_t *count, int64_t *capacity);

/* Copy the nelem == count stored elements to host arrays (any may be NULL), as
 * save / save_pimn_sf read d_hypersf, d_sf_txyz, d_pi_sf (pyvisc/ideal.py:301-313,
 * pyvisc/visc.py:401-419). */
int cv_hypersf_get(cv_handle *h, float *sf, float *txyz, float *pisf, float *bksf, int64_t nelem);

/* count <- 0 and drop the snapshot. */
int cv_hypersf_reset(cv_handle *h);

/* ---- event batches -----------------------------------------------------------------
 * An ensemble of events with their own grids, stepped in lockstep by one persistent
 * launch per RK stage (the reference runs one CLVisc per event, pyvisc/ebe.py; this is
 * the batched replacement of that loop).  All events share the physics parameters of
 * `base` (dims replace nx, ny, nz), the EOS, TAU0 and DT; each keeps its own fields, tile
 * and march chunk.  A step is a captured graph (two counter resets, two stage launches,
 * a step-counter update) replayed per step; results are identical to stepping each
 * event alone with cv_visc_steps_async. */
typedef struct cv_batch cv_batch;

/* dims: int32[3 * nevents] = (nx, ny, nz) per event.  Every event is a viscous handle
 * (flags as in base; the test flags are rejected; CV_FLAG_FP64_ROOT selects the stage
 * kernels with the FP64 root branch). */
int cv_batch_create(const cv_params *base, const cv_eos *eos, int nevents, const int32_t *dims,
                    int device, cv_batch **out);
int cv_batch_destroy(cv_batch *b);

/* The single-event handle of one event, for the per-event entry points (fields, initial
 * state, max e, tile and chunk, hypersurface).  Valid until cv_batch_destroy; queued
 * batch work has finished when it returns. */
int cv_batch_handle(cv_batch *b, int event, cv_handle **out);

/* 0 leaves an event out of the following steps (e.g. after its freeze-out). */
int cv_batch_set_active(cv_batch *b, const int32_t *active);

/* nsteps RK2 steps of every active event from loop index `loop` (cv_batch_info; it
 * starts at 0 and advances with the steps, cv_batch_set_loop moves it). */
int cv_batch_steps_async(cv_batch *b, int64_t nsteps);
int cv_batch_set_loop(cv_batch *b, int64_t loop);
int cv_batch_sync(cv_batch *b);
int cv_batch_last_elapsed_ms(cv_batch *b, float *ms);
/* events, loop index of the next step, work items (tile column x x-chunk) per stage (of the
 * last build of the event list; before the first step, of the events' own chunks) */
int cv_batch_info(cv_batch *b, int32_t *nevents, int64_t *loop, int64_t *nwork);
/* Work ranges of the persistent stage (no counterpart in the reference): the event list is
 * cut into nxcd contiguous ranges with a counter each and workgroup b starts on range
 * b % nxcd, the XCD it runs on; 0 or 1 is one shared counter.  Default 8.  Results are the
 * same either way. */
int cv_batch_set_xcd_map(cv_batch *b, int nxcd);

/* ---- learned seed of the conserved-to-primitive root solve ---------------------------
 * No reference counterpart: the reference starts every root solve at |v| = 0.5
 * (helper.h:128-167).  A 4-64-64-2 MLP (tanh) maps the features (log2 T^tt / 16,
 * |T^ti| / T^tt, T^t eta / T^tt, octave of T^tt / 24) to (e / T^tt, |v|); the root solve
 * starts at the predicted |v| and keeps its bracket [0, 1] and stopping rule, so the
 * network can only change where the iteration starts.  Training and inference run on the
 * device: BF16 operands on MFMA, FP32 accumulation, FP32 master weights and Adam.
 * Parameter vector (MLP_NPARAM = 4610 floats): W1[64][4], b1[64], W2[64][64], b2[64],
 * W3[2][64], b3[2]. */
typedef struct cv_mlp cv_mlp;
int cv_mlp_nparam(void);
int cv_mlp_create(int device, uint64_t seed, cv_mlp **out);   /* Glorot-uniform weights, zero biases */
int cv_mlp_destroy(cv_mlp *m);
int cv_mlp_get_params(cv_mlp *m, float *host, int n);
int cv_mlp_set_params(cv_mlp *m, const float *host, int n);   /* also resets Adam */
/* Gradient of the loss mean_samples sum_o (y_o - t_o)^2 at the current parameters
 * (X: n x 4 features, T: n x 2 targets); no update. */
int cv_mlp_grad(cv_mlp *m, const float *X, const float *T, int64_t n, float *grad, float *loss);
/* Training set to the device; cv_mlp_train runs Adam (beta 0.9, 0.999, eps 1e-8) on
 * consecutive batches of it (wrapping), the loss of each step into loss_hist. */
int cv_mlp_set_data(cv_mlp *m, const float *X, const float *T, int64_t n);
int cv_mlp_train(cv_mlp *m, int steps, int batch, float lr, float *loss_hist);
int cv_mlp_forward(cv_mlp *m, const float *X, int64_t n, float *Y);
/* The network's input features of given states (T^tt, |T^ti|, T^t eta) as the device forms
 * them for the seeded root solve (n x 4 into X): cv_mlp_forward of these gives the seeds
 * that cv_root_eval and the seeded stages start from. */
int cv_mlp_features(cv_mlp *m, const float *T00, const float *M, const float *T03, int64_t n, float *X);
/* Root solve of given states with the handle's EOS, plain (m NULL) or seeded: e, |v|,
 * iterations. */
int cv_root_eval(cv_handle *h, cv_mlp *m, const float *T00, const float *M, const float *T03, int64_t n,
                 float *e_out, float *v_out, int32_t *iters);
/* Ideal stages of the handle (cv_ideal_stage, cv_ideal_steps_async) seed their root solve
 * with m (NULL: the reference's start), and so do the viscous stages (cv_set_root_split
 * picks the split or the in-stage form); track_iters keeps the iteration count per cell
 * of the last stage for cv_get_root_iters. */
int cv_set_root_seed(cv_handle *h, cv_mlp *m, int track_iters);
/* No counterpart in the reference (north-star extension, the learned + batched form).  The
 * batch's stages seeded by m (NULL: back to the unseeded stages).  With an octave table the
 * network runs inside the persistent stage launch (each wave seeds its 64 cells before the
 * root solve); with a row/col table the stages run split: one persistent stage launch writes
 * every event's root-solve input, then the network and the root solve run over all events
 * inside the same graph.  Cells with |T^ti| > fp64_root_vmin T^tt take the FP64 root under
 * CV_FLAG_FP64_ROOT, the others start from the seed.  Every event's own handle gets m as
 * with cv_set_root_seed, so an event stepped alone gives the same fields. */
int cv_batch_set_root_seed(cv_batch *b, cv_mlp *m);
/* Viscous stages of the tiled kernels (kinds 1, 2) in split form: the stage writes the
 * root solve's conserved input (field CV_ROOT_CONS) and k_visc_root finishes the cells
 * (update_ev from the clamp on, kernel_visc.cl:322-393), seeded by the network set with
 * cv_set_root_seed (field CV_ROOT_SEED) or from the reference's start; with
 * CV_FLAG_FP64_ROOT the fast cells take the FP64 root, seed or not.  Without a seed the
 * split step is bit-identical to the fused one.  A network set on a viscous handle without
 * one turns the split on by itself.  on = 0 with a network set runs the seed inside the
 * stage kernel (kind 2 with an octave table, no test flags; other stages stay split): each
 * wave evaluates the network on its 64 cells before the root solve, with the same seeds,
 * iteration counts and fields as the split form bit for bit; the seeds (CV_ROOT_SEED) and
 * iteration counts are written only when cv_set_root_seed tracks iterations. */
int cv_set_root_split(cv_handle *h, int on);
int cv_get_root_iters(cv_handle *h, int32_t *host, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* CLVISC_AMD_H */
