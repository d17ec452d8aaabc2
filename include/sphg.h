/*
 * sphg.h -- C ABI of the MI355X-native per-step particle loop of the two-phase
 * (water-soil) weakly compressible SPH solver.
 *
 * The reference solver is Fortran.  Its per-step loop is
 *   time_integration (declarations_sph.f90:525)
 *     -> single_step (declarations_sph.f90:1131)
 *        -> direct_find / link_list (m_particles.f90:938, :1146)
 *        -> con_density, delta_sph, shear_strain_rate, pressure, ... (m_particles.f90)
 *        -> art_visc (m_particles.f90:2504)
 *        -> direct_find_2, darcy_law, pore_water_pressure, volume fractions
 * and it operates on `type(particles)` (m_particles.f90:62-384), which is not
 * C-interoperable.  This header is the boundary that replaces those call
 * sites: a `bind(C)` Fortran wrapper (sources_amd/fortran/sphg_mod.f90) packs
 * the particle arrays of one phase into `sphg_phase_view` and calls the entry
 * points below, which run hand-written HIP kernels on one GPU.
 *
 * Conventions
 *  - Host arrays are owned by the caller and keep the reference layout:
 *    vectors are Fortran x(2,maxn), i.e. interleaved [x1,y1,x2,y2,...];
 *    scalars are length maxn; pair arrays are length max_interaction and
 *    dwdx is dwdx(2,max_interaction).  Pair indices are 1-based particle
 *    numbers, as in the reference.
 *  - The library owns device memory.  It never frees or resizes caller memory.
 *    A view stays bound until the next sphg_bind_phase on the same phase.
 *  - Every entry point returns 0 (SPHG_OK) or a negative error code; the
 *    codes mirror the reference's `stop` statements.
 *  - One host thread per context.  Calls are synchronous at the boundary:
 *    on return, every array the entry point documents as output has been
 *    copied back into the bound host view.  The exceptions are
 *    sphg_time_integration(..., sync=0) and sphg_single_step(..., sync=0),
 *    which leave the state on the device until sphg_sync().
 */
#ifndef SPHG_H
#define SPHG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define SPHG_OK                      0
#define SPHG_ERR_TOO_MANY_INTERACTIONS (-1) /* m_particles.f90:991-994, 1052-1055, 1251-1253 */
#define SPHG_ERR_OUT_OF_RANGE        (-2)  /* grid_geom, m_particles.f90:1131-1136 */
#define SPHG_ERR_BAD_KERNEL          (-3)  /* get_scale_k, m_particles.f90:932 */
#define SPHG_ERR_INVALID             (-4)  /* bad argument / unbound phase */
#define SPHG_ERR_HIP                 (-5)  /* HIP runtime error (see sphg_last_hip_error) */
#define SPHG_ERR_UNSUPPORTED         (-6)  /* switch combination outside the ported path */
#define SPHG_ERR_NONFINITE           (-7)  /* NaN/Inf detected in a checked field */
#define SPHG_ERR_OUT_OF_MEMORY       (-8)
#define SPHG_ERR_MOVED_TOO_FAR       (-9)  /* a particle moved half the neighbour-list skin between
                                              two list rebuilds: a pair may have been missed */

/* ---- phases ------------------------------------------------------------ */
#define SPHG_WATER 1   /* `parts`, itype +-2 (m_particles.f90:896-913) */
#define SPHG_SOIL  2   /* `soil`,  itype +-3 */

/* ---- precision modes --------------------------------------------------- */
#define SPHG_FP64  0   /* every operation in FP64 (the reference's real(dp)) */
#define SPHG_MIXED 1   /* pair terms in FP32, summed in FP64; FP64 leapfrog,
                          density and Drucker-Prager projection */

/* ---- field masks for sphg_sync --------------------------------------- */
#define SPHG_F_X        (1ull << 0)
#define SPHG_F_VX       (1ull << 1)
#define SPHG_F_RHO      (1ull << 2)
#define SPHG_F_P        (1ull << 3)
#define SPHG_F_C        (1ull << 4)
#define SPHG_F_VOF      (1ull << 5)
#define SPHG_F_STRESS   (1ull << 6)   /* sxx, sxy, syy */
#define SPHG_F_EPSP     (1ull << 7)   /* epsilon_p */
#define SPHG_F_DRHO     (1ull << 8)
#define SPHG_F_DVX      (1ull << 9)
#define SPHG_F_DVOF     (1ull << 10)
#define SPHG_F_DSTRESS  (1ull << 11)  /* dsxx, dsxy, dsyy, dp */
#define SPHG_F_AV       (1ull << 12)
#define SPHG_F_STRAIN   (1ull << 13)  /* txx, txy, tyy (+ vcc, wxy for soil) */
#define SPHG_F_VOF2     (1ull << 14)
#define SPHG_F_FAIL     (1ull << 15)  /* fail(), nfail */
#define SPHG_F_COUNTIAC (1ull << 16)
#define SPHG_F_MIN      (1ull << 17)  /* rho_min, v_min, vof_min */
#define SPHG_F_ALL      (~0ull)

/*
 * Simulation parameters.  Field meanings and defaults follow the reference:
 *   material  -> `type material`   (m_particles.f90:19-29), set from the
 *                steering keywords WATER PROPERTY / SOIL PROPERTY /
 *                DARCY PARAMETER / SOIL FAILURE CRITERIA
 *                (declarations_sph.f90:280-312)
 *   numerical -> `type numerical`  (m_particles.f90:31-55)
 *   switches  -> module param      (param.f90:78-103)
 *   grid box  -> `type particles`  (m_particles.f90:89-102), used by link_list
 */
typedef struct sphg_params {
    /* water: rho0, b, gamma, c, viscosity */
    double water_rho0, water_b, water_gamma, water_c, water_viscosity;
    /* soil: rho0, k (bulk modulus), E, niu, c, porosity, permeability,
       cohesion, phi (radians) */
    double soil_rho0, soil_k, soil_E, soil_niu, soil_c;
    double soil_porosity, soil_permeability, soil_cohesion, soil_phi;
    /* numerical: nnps (1 direct_find, 2 link_list), skf (1..4) */
    int32_t nnps, skf;
    double numeric_gravity;           /* numerical%gravity, used by darcy_law */
    double alpha, beta, etq;          /* artificial viscosity */
    double dd, p1, p2;                /* wall repulsion */
    double delta;                     /* delta-SPH */
    double epsilon;                   /* XSPH */
    double gravity;                   /* declarations_sph.f90:13, added to dvx(2,:) */
    /* switches (0/1 unless stated) */
    int32_t single_phase, summation_density, average_velocity, visc_artificial;
    int32_t artificial_density, soil_artificial_stress, volume_fraction;
    int32_t water_artificial_volume, volume_fraction_renorm;
    int32_t stress_integration;       /* 1 or 2 */
    int32_t yield_criterion;          /* 1 Mohr-Coulomb, 2 Drucker-Prager */
    int32_t plasticity;               /* 0..3 */
    /* link_list sorting grid */
    double x_mingeom, x_maxgeom, y_mingeom, y_maxgeom;
    int32_t maxngx, maxngy;
    double dt;
    int32_t precision;                /* SPHG_FP64 or SPHG_MIXED */
    /* param.f90:78 integrate_scheme: 1 (or 0) leapfrog; 2 Verlet
       (time_integration_for_water_by_verlet, declarations_sph.f90:889-974,
       single-phase water only, as sph.f90:51-57 dispatches it) */
    int32_t integrate_scheme;
    int32_t reserved[6];
} sphg_params;

/*
 * Host view of one phase: the subset of `type(particles)` that the per-step
 * loop reads or writes.  Pointers may be NULL for arrays the phase does not
 * have (e.g. stress for water); an entry point that needs a NULL array
 * returns SPHG_ERR_INVALID.
 */
typedef struct sphg_phase_view {
    int32_t ntotal, nvirt, maxn, max_interaction;
    int32_t material;                 /* SPHG_WATER or SPHG_SOIL */
    int32_t niac, nfail, reserved0;   /* outputs: pair count, failed points */
    double  dspp;                     /* initial particle spacing (repulsion r0) */
    int32_t *itype, *fail, *countiac, *pair_i, *pair_j;
    double *x, *vx, *mass, *hsml;
    double *rho, *p, *c, *vof, *vof2, *dvof, *vof_min;
    double *rho_min, *v_min, *drho, *dvx, *av;
    double *sxx, *syy, *sxy, *txx, *tyy, *txy, *wxy, *vcc, *epsilon_p;
    double *dsxx, *dsyy, *dsxy, *dp;
    double *w, *dwdx;
    /* soil: the stress levels of stress_integration = 2 and the rates of
       plastic_flow_rule2 (plasticity = 2; m_particles.f90:275-293).  May be
       NULL: the device then keeps its own copies, starting at zero as
       allocate_sph leaves them (declarations_sph.f90:493-507). */
    double *sxx_min, *sxy_min, *syy_min, *p_min;
    double *dsxx2, *dsxy2, *dsyy2, *dp2;
} sphg_phase_view;

typedef struct sphg_ctx sphg_ctx;

/* ---- lifetime ---------------------------------------------------------- */

/* Create a context on the current HIP device; replaces allocate_sph
   (declarations_sph.f90:338) for the device side. */
int sphg_create(const sphg_params *params, sphg_ctx **out);
void sphg_destroy(sphg_ctx *ctx);
/* Replace the parameters (e.g. dt) of a live context. */
int sphg_set_params(sphg_ctx *ctx, const sphg_params *params);
/* Bind a phase view and upload its state (x, vx, mass, hsml, itype, rho, p,
   c, vof, stress, epsilon_p and the rate arrays) to the device.  The view
   struct itself stays in use while bound: the library writes the scalar
   outputs niac and nfail into it. */
int sphg_bind_phase(sphg_ctx *ctx, int phase, sphg_phase_view *view);
/* Copy fields between device and the bound view: to_host=1 downloads. */
int sphg_sync(sphg_ctx *ctx, int phase, uint64_t field_mask, int to_host);
/* Time step counter (declarations_sph.f90:51 itimestep). */
int sphg_set_itimestep(sphg_ctx *ctx, int itimestep);
int sphg_get_itimestep(const sphg_ctx *ctx);

/* ---- neighbour search (materialised pair lists) ------------------------ */

/* direct_find(parts), m_particles.f90:938-1001: all i<j pairs with
   sqrt(dx^2+dy^2) < scale_k*(h_i+h_j)/2, in i-major / j-ascending order,
   with w, dwdx and countiac.  Writes pair_i, pair_j, w, dwdx, countiac and
   niac of the view. */
int sphg_direct_find(sphg_ctx *ctx, int phase);
/* link_list(parts), m_particles.f90:1146-1265: the same predicate with h(1),
   restricted to the cells the reference's sorting grid reaches (init_grid /
   grid_geom, m_particles.f90:1069-1143).  Same pair SET as the reference;
   pairs are written in i-major / j-ascending order. */
int sphg_link_list(sphg_ctx *ctx, int phase);
/* direct_find_2(parts, soil), m_particles.f90:1004-1062: water i x soil j
   cross pairs (including virtual x virtual and r = 0 pairs) into the WATER
   view's pair arrays; countiac of both phases. */
int sphg_direct_find_2(sphg_ctx *ctx);

/* ---- granular physics entry points (state taken from the device) ------- */

/* con_density(parts), m_particles.f90:2402-2431: drho of the phase. */
int sphg_con_density(sphg_ctx *ctx, int phase);
/* sum_density(parts), m_particles.f90:2243-2298: Shepard summation density. */
int sphg_sum_density(sphg_ctx *ctx, int phase);
/* delta_sph on the density: adds the delta-SPH term to drho
   (m_particles.f90:2658-2695 with f = rho, df = drho). */
int sphg_delta_sph_density(sphg_ctx *ctx, int phase);
/* The internal-force group of single_step for one phase
   (declarations_sph.f90:1222-1308): shear_strain_rate, velocity_divergence,
   pressure, newtonian_fluid / drucker_prager, the df/df3 momentum, plastic
   flow rule 3 and Jaumann rate.  Overwrites dvx (and for soil ds*, dp,
   epsilon_p, fail); txx/txy/tyy/vcc/wxy become valid. */
int sphg_internal_force(sphg_ctx *ctx, int phase);
/* art_visc(parts), m_particles.f90:2504-2564: adds to dvx. */
int sphg_art_visc(sphg_ctx *ctx, int phase);

/* ---- fused step -------------------------------------------------------- */

/* One right-hand-side evaluation: single_step (two-phase,
   declarations_sph.f90:1131-1435) or single_step_for_water
   (declarations_sph.f90:1439-1526) depending on params->single_phase.
   Uses the context's itimestep.  sync=1 copies every rate/diagnostic field
   back into the views. */
int sphg_single_step(sphg_ctx *ctx, int sync);
/* nsteps leapfrog steps: time_integration (declarations_sph.f90:525-791) or
   time_integration_for_water (:794-886).  Advances itimestep.  sync=1
   downloads the whole state at the end. */
int sphg_time_integration(sphg_ctx *ctx, int nsteps, int sync);

/* ---- diagnostics / measurement ---------------------------------------- */
const char *sphg_strerror(int code);
/* The HIP error behind the last SPHG_ERR_HIP of ctx; ctx NULL: of the
   context-free entry point sphg_direct_find_noalloc. */
int sphg_last_hip_error(const sphg_ctx *ctx);
/* HIP stream (hipStream_t) all work of the context is issued on. */
void *sphg_stream(sphg_ctx *ctx);
/* Per-kernel timing with HIP events on the context stream.  When enabled,
   every launch of a named kernel is bracketed by events; stats accumulate
   until reset. */
int sphg_profile_enable(sphg_ctx *ctx, int enable);
int sphg_profile_reset(sphg_ctx *ctx);
/* Fills launches and total milliseconds for kernel `name`; returns
   SPHG_ERR_INVALID if the name is unknown. */
int sphg_profile_get(sphg_ctx *ctx, const char *name, int64_t *launches, double *total_ms);
/* Number of same-phase / cross pairs found by the most recent fused step
   (each pair once). */
int sphg_pair_counts(sphg_ctx *ctx, int64_t *water_pairs, int64_t *soil_pairs, int64_t *cross_pairs);
/* Neighbour lists of the FP64 two-phase step (no counterpart in the
   reference, whose pair search runs every step): rebuilds so far, spells of
   fast motion with the lists off, and the skin (length).  A rebuild comes
   every SPHG_NL_EVERY steps (default 64; a batch captures that period) and
   when a particle has moved a quarter of the skin or would pass 0.45 of it
   within four steps at its speed; rebuilds within 8 steps of each other turn
   the lists off for 256 steps (not in a batch).  The skin is SPHG_NL_SKIN
   (default 0.1) times the largest support radius.  SPHG_NLIST=0 turns the
   lists off (row scans every step).  All three are read at sphg_create. */
int sphg_nlist_stats(const sphg_ctx *ctx, int64_t *rebuilds, int64_t *off_spells, double *skin);

/* ---- learned multi-step surrogate ---------------------------------------
   No counterpart in the reference.  A learned advance of K steps is one
   leapfrog step of size K dt (one right-hand side, the context's precision)
   plus the correction of velocity and density towards K steps of size dt that
   a message-passing network on the same-phase neighbour graph (BF16 MFMA,
   FP32 accumulation) predicts from the state the advance starts from.  It is
   trained on samples that the plain FP64 step generates, and the plain step
   judges it: sphg_gnn_rollout runs the plain K steps beside every
   judge_every-th learned advance, measures the learned density's deviation
   from them (relative, max over real particles) and goes on from the plain
   result; the judge always runs in FP64.  After every other learned advance
   a sanity test on the device (finite, inside the grid) restores the last
   judged state when it fails, and the host, which reads its flag one advance
   later, covers the span with plain steps.  Two-phase contexts only
   (SPHG_ERR_UNSUPPORTED otherwise). */
typedef struct sphg_gnn sphg_gnn;
int sphg_gnn_nparams(void);
/* Network for K = ksteps steps with room for max_samples training samples;
   weights Xavier-uniform from SplitMix64(seed). */
int sphg_gnn_create(sphg_ctx *ctx, int ksteps, int max_samples, uint64_t seed, sphg_gnn **out);
void sphg_gnn_destroy(sphg_gnn *gnn);
/* Copies the FP32 master weights (sphg_gnn_nparams floats) to (to_host=1) or
   from (to_host=0) host memory. */
int sphg_gnn_weights(sphg_gnn *gnn, float *host, int to_host);
int sphg_gnn_num_samples(const sphg_gnn *gnn);
/* One training sample: graph and features of the current state, the coarse
   step from it, then K plain FP64 steps from it (the state advances along
   these); targets = plain result - coarse result. */
int sphg_gnn_record(sphg_gnn *gnn);
/* iters Adam steps (one sample each, cycling); loss_out: mean normalised
   loss over the last pass.  Then the gain of each correction channel is
   fitted (least squares of the prediction against the targets over every
   sample, within [0, 1.5]): the applied correction is prediction x gain. */
int sphg_gnn_train(sphg_gnn *gnn, int iters, float lr, float *loss_out);
/* The fitted gains of the correction channels (vx, vy, rho, spare). */
int sphg_gnn_gains(const sphg_gnn *gnn, float *gains);
/* One learned advance of K steps (use_net=0: the coarse step alone). */
int sphg_gnn_step(sphg_gnn *gnn, int use_net);
/* nlearned learned advances with the plain step as judge and fallback:
   nfallback: judged advances over tol plus advances replaced by plain steps
   after a failed sanity test or a failed check of the steps themselves (a
   particle past half the neighbour-list skin or out of the domain: such a
   span goes back to the last good state, as after a failed sanity test);
   max_err / rms_err:
   largest and root-mean-square relative density deviation of the judged
   learned advances from the plain steps (real particles). */
int sphg_gnn_rollout(sphg_gnn *gnn, int nlearned, int judge_every, double tol, int use_net, int *nfallback,
                     double *max_err, double *rms_err);
/* Nodes of a recorded sample (-1 if out of range). */
int sphg_gnn_sample_size(const sphg_gnn *gnn, int sample);
/* Inputs, outputs and weight gradient of one sample (for gradient checks):
   any pointer may be NULL.  grad is d(sum_w (y - t/out)^2)/dW. */
int sphg_gnn_debug(sphg_gnn *gnn, int sample, float *nf, float *raw, int *nb, int *nd, float *tgt, float *tw,
                   float *y, float *grad, float *loss, float *out_scale);

/* ---- batched scenarios -------------------------------------------------
   No counterpart in the reference (one case per process there).  Many
   independent two-phase contexts advance together by replays of one hipGraph
   per period (the neighbour-list rebuild period, 64 steps; without lists the
   reorder period, 16).  Staged form (the default): every step of the period,
   the rebuild included, is one launch per kernel for all contexts (a tile
   table maps each workgroup to a context and a block of it); a kernel without
   a batched form runs context by context, and a step whose kernels differ
   between contexts (e.g. smoothing lengths uniform in some and not in others)
   runs context by context over a few streams.  Results are bit-identical to
   advancing each context alone.  While batched, a context keeps its device
   buffers in place (graph mode); use it only through the batch until
   sphg_batch_destroy.  Per-launch profiling must be off.  SPHG_BATCH_STAGED=0
   in the environment at sphg_batch_create captures every step context by
   context over the streams instead. */
typedef struct sphg_batch sphg_batch;
int sphg_batch_create(sphg_ctx **ctxs, int n, sphg_batch **out);
void sphg_batch_destroy(sphg_batch *batch);
/* nsteps leapfrog steps of every context (eager up to the next period
   boundary, graph replays, eager remainder); checks every context's flags. */
int sphg_batch_run(sphg_batch *batch, int nsteps);
/* Number of streams the contexts are dealt over (drops the captured graph). */
int sphg_batch_set_streams(sphg_batch *batch, int nstreams);
/* Autotuner: times trial_periods replays with 1, 2, 4, ... streams (up to
   one per context) and the same steps launched eagerly on the contexts' own
   streams, and keeps the fastest (best_streams = 0: eager).  The contexts
   advance during tuning; numerics do not depend on the choice. */
int sphg_batch_tune(sphg_batch *batch, int trial_periods, int *best_streams, double *best_ms_per_step);
/* 1: the batch's captured steps are batched launches (staged form); 0:
   context by context over the streams; -1: batch is NULL. */
int sphg_batch_staged(const sphg_batch *batch);

/* ---- pair search without allocation -------------------------------------
   No counterpart in the reference.  For callers that cannot allocate device
   memory: code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc.  The pair search of sphg_direct_find (direct_find,
   m_particles.f90:938-1001: all i<j pairs with sqrt(dx^2+dy^2) <
   scale_k*(h_i+h_j)/2, i-major / j-ascending, 1-based, with w, dwdx and
   countiac) on caller-owned DEVICE arrays: x is x(2,n) interleaved, hsml has
   n entries; pair_i, pair_j, w (max_interaction) and dwdx
   (2*max_interaction) receive the pairs, countiac (n) the partner counts.
   niac (host) receives the pair count, also when it exceeds max_interaction
   (then SPHG_ERR_TOO_MANY_INTERACTIONS and no pair is written).  The grid box
   of params is the geometry box; a particle outside it gives
   SPHG_ERR_OUT_OF_RANGE.  stream: a hipStream_t (NULL: the null stream); the
   call returns after its work is complete.
   The library allocates nothing for it: its scratch is one static device
   array per GPU (8.0 MiB, held until the process ends), which
   limits n to SPHG_NOALLOC_MAX_N (larger n: SPHG_ERR_INVALID) and the cell
   grid to SPHG_NOALLOC_MAX_CELLS cells (coarser cells are used above that;
   the pairs do not change).  Every call on one GPU uses that GPU's scratch:
   calls that overlap on one GPU race. */
#define SPHG_NOALLOC_MAX_N 131072
#define SPHG_NOALLOC_MAX_CELLS 262144
int sphg_direct_find_noalloc(const sphg_params *params, int n, const double *x, const double *hsml,
                             int max_interaction, int32_t *pair_i, int32_t *pair_j, double *w, double *dwdx,
                             int32_t *countiac, int32_t *niac, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SPHG_H */
