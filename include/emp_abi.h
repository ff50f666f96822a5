/*
 * emp_abi.h -- C ABI of the MI355X (gfx950) Keplerian-RV likelihood and
 * parallel-tempering engine.
 *
 * This is the drop-in boundary for the reference's per-walker callables and
 * its sampler engine.  In the reference those live in a Python script that
 * EMPEROR writes at run time:
 *   my_model(theta) -> (model, err2)   emp/emp_model.py:706-781
 *   my_likelihood(theta) -> float      support/likelihoods/00.like:3-5
 *   my_prior(theta) -> float           emp/emp.py:182-254
 *   reddemcee.PTSampler(...)           emp/emp.py:576-593
 *   sampler.run_mcmc(p1, nsweeps, nsteps)  support/endit_reddemcee.scr:3
 *   sampler.get_chain/get_log_like/get_log_prob/get_tsw/get_smd
 *                                      emp/emp.py:1401-1429, 1985-1990
 * (emp/ = src/astroemperor/).  Each entry point below names the callable it
 * replaces.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative EMP_E* code on
 *    failure; emp_last_error() then describes the failure (thread-local).
 *  - No exceptions cross the boundary; the caller owns every buffer.
 *  - theta buffers hold the FREE parameters (fixed ones are re-inserted
 *    on device, as emp_model.py:709-711 does), row-major [B, ld], ld >= ndim.
 *  - By default pointers are device pointers and work is enqueued on
 *    `stream` (a hipStream_t, NULL = default stream) without synchronising.
 *    With EMP_HOST_POINTERS the library copies through its own staging
 *    buffers and returns after the results are on the host.
 *  - A likelihood that is not finite (the reference would return NaN, e.g.
 *    for e >= 1) is written as -inf and counted (emp_model_nonfinite_count).
 *  - A handle must not be used from two threads at once.
 */
#ifndef EMP_ABI_H
#define EMP_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMP_ABI_VERSION 3

/* Error codes. */
#define EMP_OK 0
#define EMP_EINVAL (-1)     /* bad argument or descriptor */
#define EMP_EHIP (-2)       /* HIP runtime error */
#define EMP_ENOMEM (-3)     /* device allocation failed */
#define EMP_EUNSUP (-4)     /* model feature outside what the kernels implement */
#define EMP_EIO (-5)        /* checkpoint file error */
#define EMP_EFAULT_INJECTED (-6) /* failure forced by EMP_FAULT_AT (testing) */

/* Flags. */
#define EMP_HOST_POINTERS 1
#define EMP_GRAD_MIXED 2    /* emp_lnl_grad_batch: the mixed path's Kepler solve (EMP_MIXED accuracy) */

/* Precision of the likelihood path. */
#define EMP_FP64 0          /* FP64 arithmetic (the Kepler starter's roots and reciprocals at FP32 accuracy, which
                                the FP64 fifth-order correction removes): the drop-in path */
#define EMP_MIXED 1         /* FP32 Kepler iterations + FP64 phase, final step, accumulation */

/* Prior kinds (support/priors/<Kind>.prior). */
#define EMP_PRIOR_NONE (-1)
#define EMP_PRIOR_UNIFORM 0   /* a0 = log(1/(hi-lo))                        Uniform.prior:1-5 */
#define EMP_PRIOR_NORMAL 1    /* a0 = mu, a1 = s, a2 = logZ, a3 = log(s*sqrt(2*pi))  Normal.prior:1-9 */
#define EMP_PRIOR_JEFFREYS 2  /* a0 = log(1/(hi-lo)) (the reference's Jeffreys is uniform) Jeffreys.prior:1-6 */
#define EMP_PRIOR_BETA 3      /* a0 = a, a1 = b, a2 = ln B(a, b)            Beta.prior:1-6 */
#define EMP_PRIOR_FIXED 4     /* 0                                          Fixed.prior:1-2 */

/* Semantics of per-instrument moving average (support/models/moav00.model). */
#define EMP_SEM_REFERENCE 0   /* per-instrument MA is a no-op, as the reference computes it */
#define EMP_SEM_INTENDED 1    /* moav01's recurrence within each instrument's time series */

typedef struct EmpModel EmpModel;
typedef struct EmpPT EmpPT;

typedef struct EmpPriorTerm {
    int32_t kind;      /* EMP_PRIOR_* */
    int32_t block;     /* block index: the prior returns -inf after the first block whose sum is -inf */
    double lo, hi;     /* inclusive support */
    double a0, a1, a2, a3;
} EmpPriorTerm;

/* Model description: the contents of the generated script's data loader
 * (emp/emp_model.py:406-433), model writer (706-781) and prior writer
 * (emp/emp.py:182-254), in flat form. */
typedef struct EmpModelDesc {
    int32_t abi_version;        /* EMP_ABI_VERSION */
    int32_t n_epochs;           /* N = ndat */
    int32_t n_ins;              /* I: instruments (Flag 1..I) */
    int32_t n_planets;          /* k Keplerian blocks */
    int32_t parameterisation;   /* 0,1,2,3,4,6,7 (block_repo.py:23-31) */
    int32_t use_c;              /* 1: clib/ckep0*.model variant (no small-e guard) */
    int32_t acc_order;          /* Acceleration block order, 0 = none (acc.model) */
    int32_t has_jitter;         /* Jitter block present (jitter00.model) */
    int32_t ma_order;           /* MOAV order q, 0 = none */
    int32_t ma_global;          /* 1: moav01.model (one set), 0: moav00.model (per instrument) */
    int32_t semantics;          /* EMP_SEM_* */
    int32_t ndim_full;          /* length of the full parameter vector */
    int32_t ndim;               /* free parameters */
    int32_t n_blocks;           /* number of parameter blocks */
    int32_t off_kep, off_acc, off_offset, off_jitter, off_ma;  /* block starts in the full vector */
    int32_t off_sinusoid;       /* Sinusoid block (P, A, phase) start, -1 = none (sinusoid00.model) */
    int32_t off_magnetic;       /* MagneticCycle block (P, A1, A2, phase1, phase2) start, -1 = none
                                   (magneticcycle00.model).  Both terms are added to the model
                                   after the moving average, as their blocks follow MOAV. */
    int32_t off_sai;            /* StellarActivity block start (one coefficient per activity
                                   column, sai00.model), -1 = none; also added after the MA */
    int32_t n_sai;              /* activity columns, sum of sai_counts (<= 32) */
    const double* t;            /* [N] X_ = BJD - common_t, time-sorted (host) */
    const double* y;            /* [N] Y_ */
    const double* yerr;         /* [N] YERR_ */
    const int32_t* ins;         /* [N] instrument index 0..I-1 (Flag - 1) */
    const int32_t* free_index;  /* [ndim] full index of each free parameter (C_) */
    const double* fixed_value;  /* [ndim_full] values of fixed parameters (mod_fixed_) */
    const EmpPriorTerm* prior;  /* [ndim_full] in block order (kind ignored for fixed) */
    const EmpPriorTerm* ecc_prior; /* [n_planets] prior on e = S^2 + C^2 (EMP_PRIOR_NONE if absent) */
    const double* sai;          /* [N][n_sai] activity columns in the epochs' order, instrument-major
                                   columns; zero at other instruments' epochs (null if n_sai = 0) */
    const int32_t* sai_counts;  /* [n_ins] activity columns per instrument (null if n_sai = 0) */
} EmpModelDesc;

/* Replaces: the generated module top level (data, masks, constants).
 * Copies the data to device memory of the current HIP device. */
int emp_model_create(const EmpModelDesc* desc, EmpModel** out);
void emp_model_destroy(EmpModel* model);

/* Replaces: my_likelihood(theta) (support/likelihoods/00.like:3-5), batched.
 * lnl[b] = lnL(theta[b, :ndim]). precision: EMP_FP64 or EMP_MIXED. */
int emp_lnl_batch(EmpModel* model, const double* theta, int64_t B, int64_t ld,
                  double* lnl, int32_t precision, int32_t flags, void* stream);

/* Replaces: my_prior(theta) (emp/emp.py:182-254), batched. */
int emp_lnprior_batch(EmpModel* model, const double* theta, int64_t B, int64_t ld,
                      double* lnp, int32_t flags, void* stream);

/* Replaces: my_model(theta) -> (model0, err20) (emp/emp_model.py:706-781).
 * Host pointers; one theta; FP64; outputs in the data's time order. */
int emp_model_eval(EmpModel* model, const double* theta, double* model_out, double* err2_out);

/* No reference counterpart: lnL and dlnL/dtheta[free]
 * for a batch, by a hand-written adjoint kernel (FP64; with EMP_GRAD_MIXED the
 * Kepler solve of the mixed path, FP32 start and one FP64 step). grad is [B, ndim].
 * The two-pass adjoint kernel (emp_model_set_grad_kernel: MA orders above 1,
 * series shorter than 1024 epochs by default) keeps, with an active MA term,
 * a device scratch of B x n_epochs doubles in the model (grown on demand,
 * freed with the model); the single-pass kernel needs none. */
int emp_lnl_grad_batch(EmpModel* model, const double* theta, int64_t B, int64_t ld,
                       double* lnl, double* grad, int32_t flags, void* stream);

/* No reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): emp_lnl_grad_batch with device pointers that allocates nothing.
 * The two-pass kernel's moving-average scratch is a static device array of
 * EMP_NOALLOC_SCRATCH_DOUBLES doubles (64 MiB) that each device the library
 * runs on holds until the process ends; a batch needing more (with an active
 * MA term in the two-pass kernel, B x n_epochs doubles) is refused with
 * EMP_EINVAL.  The single-pass kernel uses no scratch.  Same results as
 * emp_lnl_grad_batch, bit for bit.  flags: 0 or EMP_GRAD_MIXED
 * (EMP_HOST_POINTERS is refused).  Call with the model's device current.
 * Every call on one device shares that device's array, and each launch
 * writes every element it reads, so nothing needs clearing between calls;
 * calls that overlap on one device (other streams or threads) race. */
#define EMP_NOALLOC_SCRATCH_DOUBLES ((int64_t)1 << 23)
int emp_lnl_grad_batch_noalloc(EmpModel* model, const double* theta, int64_t B, int64_t ld,
                               double* lnl, double* grad, int32_t flags, void* stream);

/* Launch shape: waves per parameter vector (1, 2, 4 or 8); 0 = automatic. */
int emp_model_set_waves_per_theta(EmpModel* model, int32_t waves_per_theta);

/* Adjoint kernel of emp_lnl_grad_batch (and of the gradient move) for this
 * model.  No counterpart in the reference (its gradients, where a caller
 * wants them, are finite differences of my_likelihood); the two kernels
 * give the same lnL and gradients to rounding.
 *   EMP_GRAD_KERNEL_AUTO: single-pass when the MA order is at most 1 and the
 *     series has at least 1024 epochs, else two-pass;
 *   EMP_GRAD_KERNEL_TWO_PASS: likelihood with the MA residuals stored
 *     (B x N doubles of scratch), then the epochs in reverse;
 *   EMP_GRAD_KERNEL_SINGLE_PASS: one pass, no scratch; EMP_EUNSUP for MA
 *     orders above 1.
 * A gradient move enabled before a change keeps the scratch it allocated:
 * switch before emp_pt_set_moves. */
#define EMP_GRAD_KERNEL_AUTO 0
#define EMP_GRAD_KERNEL_TWO_PASS 1
#define EMP_GRAD_KERNEL_SINGLE_PASS 2
int emp_model_set_grad_kernel(EmpModel* model, int32_t which);
int emp_model_get_waves_per_theta(EmpModel* model, int32_t precision, int32_t* out);

/* Number of non-finite likelihoods mapped to -inf since creation. */
int emp_model_nonfinite_count(EmpModel* model, int64_t* out);

/* ---------------------------------------------------------------------------
 * Parallel tempering (replaces reddemcee.PTSampler + run_mcmc).
 * ------------------------------------------------------------------------ */
typedef struct EmpPTDesc {
    int32_t ntemps;             /* T */
    int32_t nwalkers;           /* W (even) */
    int32_t nsteps;             /* stretch steps per sweep (engine_config['setup'][3]) */
    int32_t precision;          /* EMP_FP64 / EMP_MIXED */
    double a;                   /* stretch scale (emcee StretchMove default 2.0) */
    uint64_t seed;              /* Philox key */
    int32_t adapt_mode;         /* engine_config['adapt_mode'] (emp/emp.py:2378, 806): -1 frozen
                                 * ladder; 0 equal swap acceptance (Vousden et al. 2016); 1 equal
                                 * mean |log swap ratio|; 2 specific heat (equal thermodynamic
                                 * length, tests/02_51peg_advanced.ipynb "adapt_mode = 2  #
                                 * Specific Heat").  Modes 1 and 2 are this library's own
                                 * definitions: reddemcee's are not in the reference. */
    int32_t record;             /* 1: keep chain/lnL/lnp history on device */
    double adapt_tau;           /* engine_config['adapt_tau'] */
    double adapt_nu;            /* engine_config['adapt_nu'] */
    int64_t max_sweeps;         /* history capacity (sweeps; nsteps rows each) when record == 1 */
} EmpPTDesc;

/* What emp_pt_get returns.  The chain history has one row per stretch step
 * (R = sweeps x nsteps rows, as reddemcee's backend counts iterations): the
 * state after each step, a sweep's last row after its swap round. */
#define EMP_PT_CHAIN 0      /* [R, T, W, D] positions (history) */
#define EMP_PT_LNL 1        /* [R, T, W] */
#define EMP_PT_LNP 2        /* [R, T, W] log prior */
#define EMP_PT_BETAS 3      /* [S, T] ladder after each sweep */
#define EMP_PT_TSW 4        /* [S, T-1] swap acceptance per adjacent pair per sweep */
#define EMP_PT_SMD 5        /* [S, T-1] mean |log swap ratio| per pair per sweep */
#define EMP_PT_ACCEPTED 6   /* [T, W] accepted moves (int64) */
#define EMP_PT_STATE_X 7    /* [T, W, D] current positions */
#define EMP_PT_STATE_LNL 8  /* [T, W] */
#define EMP_PT_STATE_LNP 9  /* [T, W] */
#define EMP_PT_STATE_BETAS 10 /* [T] */
#define EMP_PT_NSWEEPS 11   /* [2] sweeps done, likelihoods evaluated (int64) */
#define EMP_PT_GRAD_STEPS 12 /* [T] gradient-move step of each temperature (needs the move enabled once) */

/* Replaces: reddemcee.PTSampler(nwalkers, ndim, my_likelihood, my_prior,
 * ntemps=, betas=, adapt_tau=, adapt_nu=, adapt_mode=) (emp/emp.py:576-593)
 * and the initial evaluation of p1 = test_init() (emp/emp.py:650-682).
 * betas: [T] host, p0: [T, W, ndim] host. */
int emp_pt_create(EmpModel* model, const EmpPTDesc* desc, const double* betas,
                  const double* p0, EmpPT** out);
void emp_pt_destroy(EmpPT* pt);

/* Replaces: sampler.run_mcmc(p1, nsweeps, nsteps) (support/endit_reddemcee.scr:3).
 * Each sweep = nsteps red/blue steps of one move (emp_pt_set_moves) at every temperature, one round
 * of adjacent swaps (hottest pair first), an optional ladder update and the
 * history record.  use_graph: capture one sweep into a hipGraph and replay it.
 * Enqueued on the engine's own stream; returns after completion. */
int emp_pt_run(EmpPT* pt, int64_t nsweeps, int32_t use_graph);

/* Replaces: the getters (get_chain/get_log_like/get_log_prob/get_betas/
 * get_tsw/get_smd/accepted, emp/emp.py:1401-1429,1977-1990).
 * Copies into host_out (capacity in bytes); shape_out[4] receives the dims
 * (unused trailing dims = 1). */
int emp_pt_get(EmpPT* pt, int32_t what, void* host_out, int64_t capacity_bytes, int64_t* shape_out);

/* Overwrites positions (and re-evaluates lnL/lnp) and/or the ladder.
 * p: [T, W, ndim] host or NULL; betas: [T] host or NULL. */
int emp_pt_set_state(EmpPT* pt, const double* p, const double* betas);

/* Swap-round implementation.  No counterpart in the reference; the three
 * give identical results (same decisions, statistics and map) and are for
 * comparisons and tests:
 *   EMP_SWAP_AUTO: in LDS when the map, likelihoods and pre-drawn pairs fit
 *     (12 B per position and per pair, up to 62 KiB), else streamed rows for
 *     at most 1024 walkers, else global memory;
 *   EMP_SWAP_LDS, EMP_SWAP_STREAM (three temperature rows at a time in LDS),
 *   EMP_SWAP_GLOBAL: that path, EMP_EINVAL when the ensemble does not allow it. */
#define EMP_SWAP_AUTO 0
#define EMP_SWAP_LDS 1
#define EMP_SWAP_STREAM 2
#define EMP_SWAP_GLOBAL 3
int emp_pt_set_swap_path(EmpPT* pt, int32_t path);

/* Replaces: sampler.select_adjustment('00') (support/endit_freeze1.scr:11),
 * which freezes the ladder; mode -1 = frozen, 0/1/2 as EmpPTDesc.adapt_mode. */
int emp_pt_set_adaptation(EmpPT* pt, int32_t adapt_mode, double adapt_tau, double adapt_nu);

/* Moves a sweep draws from (emcee's `moves=[(move, weight), ...]`, which
 * reddemcee.PTSampler accepts and emp/emp.py:2373 leaves at None, i.e. the
 * stretch move alone).  Each sweep makes one kind of move, chosen with
 * probability weight / sum(weights) from the engine's random stream.
 *   EMP_MOVE_STRETCH:  scale = a (> 1), Goodman & Weare 2010.
 *   EMP_MOVE_GRADIENT: scale = Langevin step eps (> 0); ensemble-
 *     preconditioned Metropolis-adjusted Langevin (no reference counterpart):
 *     y = x + L ((eps^2/2) beta L^T grad lnL(x) + eps xi), L L^T the
 *     covariance of the other half of the walkers at that temperature, exact
 *     Metropolis-Hastings acceptance; lnL and gradients in FP64 (adjoint
 *     kernel).  Needs 2 x T x W x (D + 1) + T x D x (2 D + 1) doubles of
 *     device memory, plus T x W x N with the moving average. */
#define EMP_MOVE_STRETCH 0
#define EMP_MOVE_GRADIENT 1
typedef struct EmpMoveDesc {
    int32_t kind;               /* EMP_MOVE_* */
    double weight;              /* >= 0; the weights need not sum to one */
    double scale;               /* a for the stretch move, eps for the gradient move */
} EmpMoveDesc;

/* One or two moves, each kind at most once (default: stretch alone, weight 1,
 * a from EmpPTDesc).  A gradient move sets the step of every temperature to
 * its scale. */
int emp_pt_set_moves(EmpPT* pt, const EmpMoveDesc* moves, int32_t n);

/* No reference counterpart: the gradient move's step per temperature
 * (eps: [T] host, each finite and > 0), e.g. adapted during burn-in towards a
 * target acceptance and then held fixed.  Needs the gradient move enabled. */
int emp_pt_set_gradient_steps(EmpPT* pt, const double* eps);

/* Evidence ingredients for reddemcee's get_evidence_ti / get_evidence_ss
 * (called at emp/emp.py:1434-1446; the estimators run on the host): from the
 * recorded lnL of history rows [discard, rows recorded), per temperature t,
 * mean_out[t] = mean lnL, and lse_out[t] (t >= 1) = log mean exp((beta_{t-1} -
 * beta_t) lnL_t), the stepping stone to the next colder temperature under the
 * current ladder (lse_out[0] = 0), with var_out[t] its delta-method variance
 * (autocorrelation ignored).  Host arrays of T doubles.  Needs record == 1. */
int emp_pt_lnl_moments(EmpPT* pt, int64_t discard, double* mean_out, double* lse_out, double* var_out);

/* No reference counterpart (the reference cannot resume a run): binary
 * checkpoint of the full engine state, bitwise-resumable, with the recorded
 * history (ladder, tsw, smd per sweep; chain, lnL, ln prior rows when the
 * engine records).  Restore checks the engine shape (T, W, D, nsteps), the
 * sweep count against the capacity and the ladder (betas in [0, 1]).  A
 * recording engine restored from a checkpoint without chain rows holds the
 * history from the restored sweep on: getters return rows from there and
 * emp_pt_lnl_moments counts `discard` from there. */
int emp_pt_checkpoint(EmpPT* pt, const char* path);
int emp_pt_restore(EmpPT* pt, const char* path);

/* Average device time of the last emp_pt_run's sweeps in milliseconds
 * (HIP events on the engine stream), and of its half-step kernels. */
int emp_pt_last_timing(EmpPT* pt, double* ms_per_sweep, double* ms_halfstep_kernel);

/* No reference counterpart (measurement): runs ceil(n / (2 nsteps)) gradient
 * sweeps launch by launch (no graph) and returns the mean duration of the
 * first n adjoint launches over the proposals (T W / 2 vectors each), timed
 * with HIP events recorded around each launch on the engine stream.  The
 * sweeps are ordinary ones.  Needs the gradient move enabled. */
int emp_pt_time_gradstep(EmpPT* pt, int32_t n, double* ms_avg);

/* No reference counterpart (launch shape): times graph-replayed stretch
 * sweeps (8 x trials, after one untimed replay) at 1, 2, 4 and 8 waves per
 * walker on a copy of the state, restores the state (the chain is unchanged)
 * and keeps the fastest; best_ms_sweep receives its mean ms per sweep.
 * The waves per walker set the summation order of lnL, so chains run with
 * different shapes agree to rounding, not bit for bit. */
int emp_pt_autotune(EmpPT* pt, int32_t trials, int32_t* best_waves_per_theta, double* best_ms_sweep);

/* No reference counterpart (measurement): runs ceil(n / (2 nsteps)) stretch
 * sweeps launch by launch (no graph), timing the first n likelihood kernels
 * with HIP events that their launches record (kernel start and end, on the
 * engine stream), and returns their mean duration.  The sweeps are ordinary ones: they advance the chain and use
 * history capacity. */
int emp_pt_time_halfstep(EmpPT* pt, int32_t n, double* ms_avg);

/* ---------------------------------------------------------------------------
 * Batched independent systems: S separate RV models, each with its own PT
 * ensemble of T x W walkers, advanced together one sweep per launch.
 * Replaces S separate PTSampler(...).run_mcmc(...) runs (emp/emp.py:576-593,
 * support/endit_reddemcee.scr:3), one per star.  The ensembles share T, W,
 * nsteps, a and the precision; epochs, instruments, planets and ndim differ.
 * ------------------------------------------------------------------------ */
typedef struct EmpSystems EmpSystems;

typedef struct EmpSystemsDesc {
    int32_t n_systems;          /* S */
    int32_t ntemps;             /* T per system */
    int32_t nwalkers;           /* W per system (even) */
    int32_t nsteps;             /* stretch steps per sweep */
    int32_t precision;          /* EMP_FP64 / EMP_MIXED */
    int32_t adapt_mode;         /* -1 frozen, 0/1/2 as EmpPTDesc.adapt_mode */
    double a;                   /* stretch scale */
    uint64_t seed;              /* Philox key; system s draws with s mixed into the counters */
    double adapt_tau, adapt_nu;
    int32_t record_every;       /* record the coldest temperatures every k sweeps (0: never) */
    int32_t record_temps;       /* how many of the coldest temperatures are recorded (1..T) */
    int64_t max_records;        /* record capacity */
} EmpSystemsDesc;

/* Launch configuration of the batched sweep kernel, chosen by the autotuner
 * or set by the caller. */
typedef struct EmpSystemsTuning {
    int32_t waves_per_wg;       /* 8 or 16 */
    int32_t waves_per_theta;    /* 1 or 2 */
    int32_t wgs_per_cu;         /* persistent workgroups per compute unit, 1..4 */
    int32_t order;              /* 0: longest system first (LPT), 1: input order */
    double ms_per_sweep;        /* measured by the autotuner (output) */
} EmpSystemsTuning;

/* models: [S] descriptors (the same rules as emp_model_create).
 * betas: [S, T] host.  p0: host, system after system, each [T, W, ndim_s].
 * Evaluates the initial lnL / lnp of every walker. */
int emp_systems_create(const EmpModelDesc* models, const EmpSystemsDesc* desc, const double* betas,
                       const double* p0, EmpSystems** out);
void emp_systems_destroy(EmpSystems* sys);

/* nsweeps sweeps of every system; one kernel launch per sweep (captured once
 * into a hipGraph when use_graph != 0).  Returns after completion. */
int emp_systems_run(EmpSystems* sys, int64_t nsweeps, int32_t use_graph);

/* Times every launch configuration that fits for trial_sweeps sweeps each
 * (restoring the state after each trial, so the chains are unchanged), keeps
 * the fastest and reports it in *best (may be NULL). */
int emp_systems_autotune(EmpSystems* sys, int32_t trial_sweeps, EmpSystemsTuning* best);
int emp_systems_set_tuning(EmpSystems* sys, const EmpSystemsTuning* tuning);
int emp_systems_get_tuning(EmpSystems* sys, EmpSystemsTuning* tuning);

/* Getters.  system = s selects one system; system = -1 returns all of them
 * (system after system for the ragged EMP_PT_STATE_X and EMP_PT_CHAIN).
 *   EMP_PT_STATE_X      [T, W, D_s]
 *   EMP_PT_STATE_LNL/LNP, EMP_PT_ACCEPTED (int64)   [S?, T, W]
 *   EMP_PT_STATE_BETAS  [S?, T]
 *   EMP_PT_TSW, EMP_PT_SMD  [S?, T-1] means over the sweeps run so far
 *   EMP_PT_CHAIN        [R, record_temps, W, D_s] (system >= 0 only)
 *   EMP_PT_LNL/LNP      [R, S?, record_temps, W]
 *   EMP_PT_NSWEEPS      [2] sweeps done, likelihoods evaluated (int64) */
int emp_systems_get(EmpSystems* sys, int32_t what, int32_t system, void* host_out, int64_t capacity_bytes,
                    int64_t* shape_out);

/* Average device time per sweep of the last emp_systems_run (ms). */
int emp_systems_last_timing(EmpSystems* sys, double* ms_per_sweep);

/* ---------------------------------------------------------------------------
 * Misc.
 * ------------------------------------------------------------------------ */
/* No reference counterpart (test hook, needs no device): the engine's
 * Philox-4x32-10 block function, for known-answer tests of the RNG. */
int emp_philox4x32_10(const uint32_t* ctr4, const uint32_t* key2, uint32_t* out4);

const char* emp_last_error(void);
int emp_abi_version(void);
/* Device properties used by the roofline report: compute units, max clock (kHz), gcnArch name. */
int emp_device_info(int32_t* n_cu, int32_t* clock_khz, char* arch_name, int32_t arch_cap);

#ifdef __cplusplus
}
#endif

#endif /* EMP_ABI_H */
