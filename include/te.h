/* ThunderEngine MI355X large step — C ABI.
 *
 * The drop-in boundary of the CM1 dynamics large step.  Two ways in:
 *
 *  1. Program interface: executes a planned command tape, the job of the
 *     reference's WebGPU replayer (web/replay.js:36-396, class Replayer).  A
 *     program is the replayer's op list after plan(): one flat word memory M
 *     and, per dispatch, a kernel name and a 64-word argument slot
 *     (A[0..2] thread grid, A[4+n] word offset of parameter n, replay.js:7-11,
 *     258-288); "_copy" ops move A[0] words from A[4] to A[5]
 *     (replay.js:403-406).
 *  2. Config interface: builds the program itself from the namelist flags of
 *     the hot path (the reference's buildRun()/TimeStep.encode,
 *     web/core/driver.js:23-75 and web/core/timestep.js:797-2201), lays out the
 *     FieldArena (web/core/grid.js:322-346) and fills the t=0 state.
 *
 * Conventions: every call returns TE_OK (0) or a negative status; the message
 * of the last failure is te_last_error(ctx).  The caller owns host buffers, the
 * context owns device memory and one HIP stream; calls are stream-ordered and
 * not thread-safe on one context.  Nothing here takes torch types.
 */
#ifndef TE_H
#define TE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TE_OK 0
#define TE_ERR_ARG (-1)
#define TE_ERR_HIP (-2)
#define TE_ERR_UNSUPPORTED (-3)
#define TE_ERR_STATE (-4)

typedef struct te_ctx te_ctx;

/* ---------------------------------------------------------------- programs */

/* Replayer(device, tape, kernels, sources) + plan() + build()
 * (web/replay.js:36-66,180-209): allocates mem_words words of device memory
 * (zeroed) and binds the op list.  names[i] is a reference kernel name or
 * "_copy"; args holds nops*64 words. */
int te_program_create(uint64_t mem_words, int32_t nops, const char* const* names, const uint32_t* args,
                      te_ctx** out);

/* Replayer.upload / writeWords (web/replay.js:216-229,301): host -> M. */
int te_upload_words(te_ctx* ctx, uint64_t word_off, const uint32_t* src, uint64_t nwords);

/* Replayer.read (web/replay.js:368-383): M -> host (synchronous). */
int te_download_words(te_ctx* ctx, uint64_t word_off, uint32_t* dst, uint64_t nwords);

/* Replayer.encodeSteps(enc, n, false, lastOp) (web/replay.js:335-366): runs ops
 * [first_op, last_op] of the program nsteps times (last_op < 0: to the end).
 * Full steps replay a captured HIP graph. */
int te_run(te_ctx* ctx, int32_t nsteps, int32_t first_op, int32_t last_op);

/* Runs nsteps full steps eagerly with a HIP event after every op on the
 * context's stream and adds each op's device duration (ms) into op_ms
 * (te_op_count floats, cleared first).  Profiling aid: no counterpart in the
 * reference, whose replayer only times whole steps. */
int te_profile_ops(te_ctx* ctx, int32_t nsteps, float* op_ms);

/* Waits for the context's stream. */
int te_sync(te_ctx* ctx);

/* region_hash of the reference (web/replay.js:408-418) over nwords of M. */
int te_region_hash(te_ctx* ctx, uint64_t word_off, uint64_t nwords, uint32_t* out);

/* word-wise FNV-1a (web/view.js:1064-1068) over nwords of M. */
int te_fnv1a(te_ctx* ctx, uint64_t word_off, uint64_t nwords, uint32_t* out);

/* region_hash (as te_region_hash) of n regions of a caller's device memory,
 * allocating nothing.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc) and have no context.  d_mem is a device
 * pointer on the current device; region r is counts[r] words from word
 * offs[r] (host arrays), its hash goes to out[r] (host).  The per-region
 * accumulators live in a device array of the library of
 * TE_NOALLOC_MAX_REGIONS words per device, cleared by a kernel on every call;
 * n above that limit returns TE_ERR_ARG, as do missing pointers and ranges
 * whose byte addresses overflow (the caller keeps every region inside its
 * buffer: the device memory's size is not known here).  Runs on the device's
 * null stream and waits for it.  Calls that overlap on one device share the
 * accumulators and race. */
#define TE_NOALLOC_MAX_REGIONS 256
int te_region_hashes_noalloc(const uint32_t* d_mem, const uint64_t* offs, const uint64_t* counts, int32_t n,
                             uint32_t* out);

/* Total words of M, number of ops, and op i as (name, 64 argument words). */
uint64_t te_memory_words(const te_ctx* ctx);
int32_t te_op_count(const te_ctx* ctx);
int te_op_get(const te_ctx* ctx, int32_t i, const char** name, uint32_t* args64);

/* Kernels the library implements (the pipelines of data/kernels/kernels.json
 * that belong to the large step, plus "_copy"): count and name i. */
int32_t te_kernel_count(void);
const char* te_kernel_name(int32_t i);

/* Self-test of the kernels' shortcuts (no counterpart in the reference):
 * over every binary32 input x, on device 0, counts the x where the shortcut
 * differs in any bit from what it stands for (mismatch[f]), the first such x
 * (first[f], 0xFFFFFFFF if none), and the x where the shortcut defers to it
 * (slow[f]); f = 0 log, 1 exp, 2 pow(x, cp/rd), 3 pow(x, rd/cv),
 * 4 pow(x, 1/3) (binary64 evaluations against the double-float routines the
 * kernels restate), 5 the recurrences' reciprocal against 1.0f / x. */
int te_selftest_math(uint64_t mismatch[6], uint32_t first[6], uint64_t slow[6]);

/* Device duration of the last te_run in milliseconds (HIP events on the
 * context's stream). */
int te_last_run_ms(te_ctx* ctx, float* ms);
/* Launch-shape autotuner.  No counterpart in the reference (WebGPU fixes the
 * workgroup size per pipeline).  For every dispatch on the generic 3-D path
 * (not the column sweeps or flat copies; fused advection among the shapes
 * that keep one row of cells per wave) it times candidate block shapes and
 * the (i, j)-plane numbering inside `reps` whole eager steps
 * (an event after each launch, every trial from the same memory), keeps each
 * dispatch's fastest, and then keeps that set only if `reps` whole steps run
 * at least 0.5% faster with it than with the default shapes (three
 * interleaved comparisons), or always when `force` is nonzero.  The memory is restored and a captured step
 * graph captured again.  A shape only changes which thread computes which
 * point, never the arithmetic.  te_tuned_ops: dispatches left on a tuned
 * shape (0 when the comparison rejected the set); te_autotune_ms: mean eager
 * step time of the comparison, default and tuned shapes. */
int te_autotune(te_ctx* ctx, int32_t reps, int32_t force);
int32_t te_tuned_ops(const te_ctx* ctx);
int te_autotune_ms(const te_ctx* ctx, float* untuned, float* tuned);

const char* te_last_error(const te_ctx* ctx);
void te_destroy(te_ctx* ctx);

/* ------------------------------------------------------------------ config */

/* Namelist subset of the hot path (web/core/config.js:113-207 field rules;
 * web/core/scenario.js:138-226 flag wiring).  Scalars are th, or th qv qc
 * when moist (qv, qc advected without microphysics). */
typedef struct te_config {
  int32_t nx, ny, nz;
  float dx, dy, dz;        /* m; dz is the nominal spacing when zf is given */
  const float* zf;         /* nz+1 face heights (m), or NULL for uniform dz */
  float dt;                /* large step (s) */
  int32_t nsound;          /* acoustic parameter: stage loops n1, n/2, n (timestep.js:73-83) */
  int32_t moist;           /* imoist */
  int32_t open_x, open_y;  /* wbc/ebc, sbc/nbc = 2 (open); 0 = periodic */
  int32_t weno_scalars;    /* advwenos = 2 */
  int32_t weno_momentum;   /* advwenov = 2 */
  int32_t tke;             /* sgsmodel = 1 (1.5-order TKE, LES) */
  int32_t explicit_vdiff;  /* doimpl = 0 */
  int32_t dissipative_heating; /* idiss = 1 */
  int32_t rayleigh;        /* irdamp = 1 */
  float rayleigh_zd;       /* zd (m) */
  float rayleigh_alpha;    /* rdalpha (1/s) */
  float alph, kdiv, smeps; /* acoustic parameters (driver.js:28) */
  float umove, vmove;      /* domain translation (m/s) */
  /* initial state (web/core/base.js WK82 sounding, quarter-circle hodograph, cos^2 bubble) */
  float th_sfc, th_trop, t_trop, z_trop, qv_pbl;  /* K, K, K, m, kg/kg */
  float hodo_radius, hodo_depth, shear_umax, shear_depth; /* m/s, m, m/s, m */
  float bubble_amp, bubble_rh, bubble_rv, bubble_zc;      /* K, m, m, m */
  float noise_sigma;       /* K, N(0, sigma) theta' noise */
  uint64_t seed;
  int32_t initial_tke;     /* init3d testcase of the t=0 TKE profile: 0 none, 3, 4, 5, 9, 11
                              (scenario.js:347-369) */
  int32_t fuse;            /* 1 (default): merged / fused launches where the step allows them
                              (same state bits; scratch fields may differ), on periodic domains one
                              launch per acoustic substep after acoustic_uv; 2: the same with the
                              acoustic substep as separate launches (the program the adjoint
                              recomputes); 0: the reference's dispatch list one for one */
  uint64_t sounding_seed;  /* perturbed soundings (ensemble members): seed of the draws below */
  float th_pert_sigma;     /* K: theta(z) += sigma * N(0,1) noise smoothed over ~1 km in z */
  float qv_pert_sigma;     /* qv(z) *= 1 + sigma * N(0,1), one draw per sounding */
  /* encoder options used to reproduce recorded reference configurations */
  int32_t species;         /* moist scalars: 3 = th qv qc, 7 = + qr qi qs qg, 11 = + nci ncs ncr ncg */
  int32_t microphysics;    /* 0 = none; 5 = Morrison call sites (encoded only, not executable) */
  int32_t mass_conservation;          /* apmasscon = 1: dry-mass correction at the end of the step */
  int32_t cache_moisture_turbulence;  /* moisture SGS fluxes computed once per step */
  int32_t n_arena_fields;             /* > 0: explicit FieldArena field list (sorted by the caller) */
  const char* const* arena_fields;
  /* horizontal stretching (namelist stretch_x, dx_inner, dx_outer, nos_x_len,
     tot_x_len and the y set; HorizontalGrid.stretched, web/core/grid.js:102-143):
     0 uniform dx; 1 spacing dx_inner over the central nos_x_len (m) growing to
     dx_outer on both sides up to tot_x_len (m); 2 the same with the growth on
     the east side only.  nx must equal the cell count those lengths imply
     (web/core/grid.js:117-120); dx stays the nominal spacing of the map
     factors dx / (local spacing). */
  int32_t stretch_x, stretch_y;
  double dx_inner, dx_outer, nos_x_len, tot_x_len;
  double dy_inner, dy_outer, nos_y_len, tot_y_len;
  /* adapt_dt = 1 (web/core/adaptive.js, web/core/driver.js:36-50): the step
     ends with the stability maxima (stability_max) and te_adaptive_start /
     te_adaptive_step choose each large step's dt and nsound from them; dt
     above is the namelist dtl; tapfrq, statfrq, rstfrq (s) are the output
     clocks whose times the steps land on (0: none).  Needs tke = 1
     (web/core/config.js:725-729). */
  int32_t adaptive_dt;
  float tapfrq, statfrq, rstfrq;
  /* the caller's sounding (buildRun(configuration, sounding), driver.js:23;
     isnd = 7 / 17, parseSounding config.js:899-901): the text of a CM1
     input_sounding file (header p_sfc mb, theta_sfc K, qv_sfc g/kg; rows z m,
     theta K, qv g/kg and with sounding_winds = 1 also u v m/s), interpolated
     to the grid (Sounding.interpolated, base.js:716-757); NULL: the analytic
     Weisman-Klemp sounding and hodograph above (the perturbed-sounding
     options apply to that one only).  te_encode / te_create keep a copy. */
  const char* input_sounding;
  int32_t sounding_winds;
} te_config;

/* Fills a config with the synthetic 64x64x40 case defaults (dx=dy=1 km,
 * dz=500 m, dt=6 s, nsound=6, moist, periodic, WENO scalars, WK82), with the
 * reference scenario's Rayleigh damping above 15 km and once-per-step
 * moisture SGS fluxes. */
void te_config_defaults(te_config* cfg);

/* TimeStep.encode (timestep.js:797-2201) without a device: grid, profiles,
 * arena layout and the op list of one large step (readable with te_op_get;
 * ops of kernels outside this path keep their names but cannot run). */
int te_encode(const te_config* cfg, te_ctx** out);

/* buildRun(conf) (driver.js:23-75): te_encode plus device memory with the
 * profiles and parameter bytes in place; the arena state is zero. */
int te_create(const te_config* cfg, te_ctx** out);

/* SupercellScenario.fill (web/core/scenario.js:234-296): t=0 state into the
 * arena and uploaded (for a te_encode context: into the host image only,
 * readable with te_download_words). */
int te_init_state(te_ctx* ctx);

/* run.encodeStep() + encodeSteps (driver.js:53): n large steps. */
int te_step(te_ctx* ctx, int32_t nsteps);

/* FieldArena.handles (grid.js:329-330): word offset of a field in M, words
 * per field slot, and the arena words count. */
int te_field_offset(const te_ctx* ctx, const char* name, uint64_t* word_off);
uint64_t te_field_stride_words(const te_ctx* ctx);
int32_t te_field_count(const te_ctx* ctx);
const char* te_field_name(const te_ctx* ctx, int32_t i);

/* Word offset of a named profile / parameter region of an encoded program
 * (c1 c2 mh mf zf rho0 rr0 rf0 th0 pi0 rth0 qv0 qc0 qi0 tauh tauf u0 v0
 * hmetric radbcw radbce radbcs radbcn sfc dims). */
int te_region_offset(const te_ctx* ctx, const char* name, uint64_t* word_off);

/* The named regions of a config context in memory order (names for
 * te_region_offset): the profile and parameter buffers a tape export lists
 * (buildRun's tape(), driver.js:63-73, one buffer per region). */
int32_t te_region_count(const te_ctx* ctx);
const char* te_region_name(const te_ctx* ctx, int32_t i);

/* Grid of a config context: nx ny nz strideJ strideK. */
int te_grid(const te_ctx* ctx, int32_t out5[5]);

/* ---- variable large step (adapt_dt = 1) */

/* TimeStep.setTimeStep (timestep.js:711-716): the step encoded anew for
 * large step dt (s) and acoustic parameter nsound, the state kept.  When the
 * dispatch list is unchanged (same nsound) only the parameter words change
 * and the step graph is kept; otherwise the new list replaces the old (and
 * launch shapes chosen by te_autotune are dropped). */
int te_set_time_step(te_ctx* ctx, float dt, int32_t nsound);

/* AdaptiveTimeStep (web/core/adaptive.js:15-109): CM1's getnewdt and
 * dtsmall (misclibs.F:543-847) and output clocks, in the reference's
 * mixed binary32 / double arithmetic.  freq[6]: tape, restart, stat,
 * domainDiag, azimAvg, hifrq frequencies (s; NaN = clock absent);
 * min_dz > 0 only with explicit vertical acoustics (psolver = 2). */
typedef struct te_clock te_clock;
typedef struct te_clock_state {
  double mtime;      /* model time (s) */
  double dbldt;      /* the next step (s, double) */
  float dt, dtlast;  /* the next and the previous step (s) */
  float cflmax, ksmax; /* last maxima, scaled to the new dt */
  int32_t nsound;    /* acoustic parameter of the next step */
  int32_t steps;     /* steps begun */
  uint32_t fired;    /* bit n: clock kind n fired this step (0 tape .. 5 hifrq) */
} te_clock_state;
te_clock* te_clock_create(double dtl, double min_dx, double min_dz, int32_t two_dimensional, int32_t ks_check,
                          const float freq[6]);
void te_clock_start(te_clock* c, float cflmax, float ksmax);  /* start(): t = 0 maxima */
void te_clock_begin_step(te_clock* c);                         /* beginStep() */
void te_clock_end_step(te_clock* c, float cflmax, float ksmax); /* endStep(): the step's maxima */
void te_clock_get(const te_clock* c, te_clock_state* out);
void te_clock_destroy(te_clock* c);

/* The driver's adaptive run (driver.js:43-49, view.js:513-550) on a context
 * created with adaptive_dt = 1.  te_adaptive_start: the stability maxima of
 * the current state (stability_max alone, encodeInitial) start the clock and
 * set the first step.  te_adaptive_step: n times beginStep, one large step,
 * read and clear the two stability words (stabilityFromWords: square roots
 * of the maxima), endStep, te_set_time_step with the clock's dt and nsound.
 * te_adaptive_clock: the clock's state. */
int te_adaptive_start(te_ctx* ctx);
int te_adaptive_step(te_ctx* ctx, int32_t nsteps);
int te_adaptive_clock(const te_ctx* ctx, te_clock_state* out);

/* ---- ensembles: members of uneven grids and perturbed soundings stepped
 * together.  No counterpart in the reference (one run per page there).  Each
 * member is a te_create context with its t=0 state; te_ensemble_step
 * replays every member's step graph on the member's own stream, step by
 * step, so the device runs members concurrently. */
typedef struct te_ensemble te_ensemble;
int te_ensemble_create(const te_config* members, int32_t n, te_ensemble** out);
int te_ensemble_step(te_ensemble* ens, int32_t nsteps);
int te_ensemble_sync(te_ensemble* ens);
int32_t te_ensemble_size(const te_ensemble* ens);
te_ctx* te_ensemble_member(te_ensemble* ens, int32_t m);  /* borrowed: valid until te_ensemble_destroy */
uint64_t te_ensemble_points(const te_ensemble* ens);        /* interior points summed over members */
const char* te_ensemble_last_error(const te_ensemble* ens);
void te_ensemble_destroy(te_ensemble* ens);

/* ---- adjoint of the large step.  No counterpart in the reference (its step
 * is forward only).  The adjoint of the reference's dispatch list: every
 * dispatch of the step has an adjoint launch over its thread grid that adds
 * J^T ybar of that dispatch to an adjoint memory laid out like M (one float
 * per word); acoustic_w_solve's adjoint solves the transposed tridiagonal
 * system per column.
 *
 * te_adjoint_create(ctx, every): ctx is a config context (fused or not); the
 *   adjoint builds the same configuration's unfused dispatch list on memory of
 *   its own (same profiles and FieldArena offsets) and a zeroed adjoint memory.
 *   Unsupported kernels (open boundaries, turbulence, microphysics) make it
 *   fail with TE_ERR_UNSUPPORTED naming the kernel.
 * te_fwd_ckpt(adj, n): runs n large steps of ctx (its fast path), copying the
 *   state (profiles + arena) to a checkpoint before every `every`-th step.
 * te_adj_upload / te_adj_download / te_adj_zero: the adjoint memory (seed it
 *   with dJ/d(state after the n steps), read dJ/d(state before them)).
 * te_bwd(adj): the reverse sweep over the n recorded steps: per checkpoint
 *   segment, the step-start states are recomputed from the checkpoint, and
 *   each step, last first, is rerun dispatch by dispatch with the regions each
 *   dispatch writes saved; then, dispatch by dispatch from the last, those
 *   regions are restored and the dispatch's adjoint launched.  The adjoint
 *   memory then holds dJ/d(state at the first recorded step).
 * te_adj_ops(adj, first, last): adjoint launches of dispatches [first, last]
 *   of the unfused list, last first, on the adjoint's memory as it stands (no
 *   restores): one dispatch's J^T for tests.
 * te_adjoint_program(adj): the unfused context (borrowed) whose memory the
 *   adjoint launches read. */
typedef struct te_adjoint te_adjoint;
int te_adjoint_create(te_ctx* ctx, int32_t every, te_adjoint** out);
te_ctx* te_adjoint_program(te_adjoint* adj);
int te_fwd_ckpt(te_adjoint* adj, int32_t nsteps);
int te_bwd(te_adjoint* adj);
int te_adj_ops(te_adjoint* adj, int32_t first_op, int32_t last_op);
int te_adj_upload(te_adjoint* adj, uint64_t word_off, const float* src, uint64_t nwords);
int te_adj_download(te_adjoint* adj, uint64_t word_off, float* dst, uint64_t nwords);
int te_adj_zero(te_adjoint* adj);
int te_adj_last_ms(te_adjoint* adj, float* fwd_ms, float* bwd_ms);
const char* te_adjoint_last_error(const te_adjoint* adj);
void te_adjoint_destroy(te_adjoint* adj);

#ifdef __cplusplus
}
#endif
#endif
