/*
 * pulsar_rocm.h — C-ABI of libpulsar_rocm.so, the MI355X (gfx950) GRAPE inner loop.
 *
 * One evaluation computes, for every problem of a plan, the ensemble-weighted
 * fidelity F and its gradient G[n_ctrl x n_t] with respect to the
 * piecewise-constant controls:
 *
 *   X_k  = c * dt_k * (A0_m + sum_j s_mj u_jk A_j)    c = -i (closed), 1 (Lindblad)
 *   U_k  = exp(X_k)                                   one per member m and slice k
 *   P_k  = U_{k-1} ... U_1,  R_k = U_n ... U_{k+1}   forward / backward chains
 *   F    = sum_m w_m F_m
 *   G_jk = sum_m w_m dF_m/du_jk
 *
 * The gradient follows the reference's first-order GRAPE formula
 * (grad_mode 0) or the exact Frechet derivative of exp (grad_mode 1).
 *
 * Replaces, on the :rocm device, the reference's
 *   compute_fidelity(sys, ctrl, tgt)          src/Physics/Objectives.jl:632-639
 *   compute_grape_gradient(sys, ctrl, tgt)    src/Physics/Gradients.jl:80-127
 *   _gate_gradient_into! / _state_gradient_into!  src/Physics/Gradients.jl:254-291, 357-399
 *   ensemble_value_and_grad batched_eval      src/Optimization/Ensemble/EnsembleObjective.jl:272-307
 *   grape_state_kernel (:cpu/:cuda branches)  src/Application/MR/GRAPEState.jl:73-85, 136-250
 *   grape_lindblad_kernel                     src/Application/MR/GRAPELindblad.jl:83-92, 96-215
 *   grape_tracking_kernel                     src/Application/MR/GRAPETracking.jl:145-263
 *   compute_grape_gradient(_powder) for MAS   src/Physics/MRPhysics.jl:155-188, 312-336
 *
 * Conventions (identical to Julia's in-memory layout, so a ccall passes
 * arrays without copies):
 *   - complex matrices are ComplexF64, column-major, interleaved (re, im):
 *     element (r, c) of a d x d matrix is at doubles [2*(r + c*d)], [2*(r + c*d)+1];
 *   - controls and gradients are Float64 [n_ctrl x n_t] column-major:
 *     u_jk at [j + k*n_ctrl]  (the reference's ControlSequence.controls view,
 *     src/Types/ControlSequence.jl:99-110);
 *   - all entry points return PR_OK (0) or a negative PR_ERR_* code; the
 *     message of the last failure on the calling thread is read with
 *     pr_last_error.
 *
 * The library never keeps a caller's host pointer beyond the call that
 * receives it. A plan owns its device memory; a plan is not re-entrant (one
 * caller at a time per plan).
 */
#ifndef PULSAR_ROCM_H
#define PULSAR_ROCM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PR_ABI_VERSION 2

/* status codes */
#define PR_OK 0
#define PR_ERR_ARG (-1)         /* invalid argument / shape / non-Hermitian input */
#define PR_ERR_DEVICE (-2)      /* no usable gfx950 device, or a HIP runtime error */
#define PR_ERR_NOMEM (-3)       /* device allocation failed */
#define PR_ERR_LAUNCH (-4)      /* kernel launch failed */
#define PR_ERR_UNSUPPORTED (-5) /* valid input outside what this build supports */
#define PR_ERR_INJECTED (-6)    /* failure forced by PULSAR_ROCM_FAULT (tests only) */

/* dynamics */
#define PR_CLOSED 0   /* Hilbert space, generator -i*H (H Hermitian) */
#define PR_LINDBLAD 1 /* Liouville space, generator L (any complex matrix, dim = N^2) */

/* objectives */
#define PR_GATE 0         /* F_m = |Tr(U_t^dag U_total)/dim|^2 (Objectives.jl:239-246); gradient of
                             _gate_gradient_into! (Gradients.jl:254-291) with the backward
                             propagators its caller builds: Q_k^dag = U_n ... U_k, slice k
                             included (Propagators.jl:375-380) */
#define PR_PAIRS_SQUARE 1 /* F_m = sum_s |<t_s|U_total|i_s>|^2 (Objectives.jl:221); gradient of the
                             MR state kernel (GRAPEState.jl:218-229): <lambda_{k+1}|A_j|psi_k> */
#define PR_PAIRS_REAL 2   /* F_m = sum_s Re <t_s|U_total|i_s>  (Objectives.jl:220), as above */
#define PR_TRACKING 3     /* F_m = sum_s sum_c w_c Re Tr(rho_c P_{n_c} |i_s><i_s| P_{n_c}^dag),
                             P_n = U_n ... U_1: density-matrix checkpoints of
                             _grape_tracking_cpu_threaded (GRAPETracking.jl:162-263);
                             its gradient is that kernel's co-state formula
                             -dt Re(i Tr(Lambda^dag [s_mj A_j, rho_{n+1}])) (:238-248).
                             Closed dynamics, dim <= 64, PR_GRAD_REFERENCE; runs FP64
                             whatever precision says */
#define PR_STATE 5        /* closed dynamics: F_m as PR_PAIRS_SQUARE; gradient of
                             _state_gradient_into! (Gradients.jl:357-399), lambda_k = Q_k psi_t
                             with slice k included: 2 Re(conj(z) (-i dt) <lambda_k|A_j|P_k psi_i>)
                             (compute_grape_gradient with a state target; the caller
                             normalises the states as Gradients.jl:369-374 does) */
#define PR_STATE_PROPAGATORS 6 /* closed dynamics, state branch of _grape_gradient_from_propagators
                                  (MRPhysics.jl:337-351): F_m as PR_PAIRS_SQUARE; gradient
                                  2 Im(conj(z) <Q_k psi_t| -i dt A_j |P_k psi_i>), states as given.
                                  PR_GRAD_REFERENCE */
#define PR_PAIRS_MODULUS 7 /* closed dynamics: F_m = sum_s |<t_s|U_total|i_s>| (Objectives.jl:222);
                              gradient of the MR state kernel with the :modulus prefactor
                              dt Im(inner / |z_s|), 0 when |z_s| < 1e-14 (Objectives.jl:356-361;
                              accepted by GRAPEState via OptControl.jl:647) */
#define PR_GATE_PROPAGATORS 4 /* F as PR_GATE; gradient of _grape_gradient_from_propagators
                                 (src/Physics/MRPhysics.jl:312-336, the MAS/DNP gate gradient):
                                 G_jk = 2 Re(conj(Phi) Tr(Q_k^dag (-i dt A_j) P_k)) / dim^2,
                                 P_k = U_{k-1}...U_1, Q_k^dag = U_n...U_k (no U_target^dag in
                                 the trace). dim <= 64, PR_GRAD_REFERENCE */

/* precision of the device path */
#define PR_FP64 0  /* every product in FP64 (FP64 MFMA) */
#define PR_MIXED 1 /* d <= 16 closed dynamics, reference gradient: the exp polynomial
                      beyond first order and the gradient walk's increments on
                      FP32 MFMA, accumulated in FP64; generators, forward chain,
                      traces and reductions FP64. Time segments with a generator
                      above the degree-6 bound, and other problems, run as PR_FP64. */

#define PR_SPLIT16 2 /* d <= 16 closed dynamics, reference gradient: as PR_MIXED with the
                        low-precision products on FP16 matrix cores (v_mfma_f32_16x16x32_f16),
                        each operand split into two FP16 parts (three products,
                        FP32-class, FP32 accumulation; the square of the exp
                        polynomial keeps the fourth, lo x lo). Time segments with a
                        generator whose 1-norm bound exceeds 0.03, and other
                        problems, run as PR_FP64. */

/* gradient formula */
#define PR_GRAD_REFERENCE 0 /* first-order GRAPE formula of Gradients.jl:274-289 */
#define PR_GRAD_EXACT 1     /* exact derivative (Frechet derivative of exp) */

/* One GRAPE problem: a system (drift + controls), an ensemble of members
 * and a target. Mirrors QuantumSystem + ControlSequence + QuantumTarget
 * (src/Types/QuantumSystem.jl:23-130, ControlSequence.jl:37-59,
 * Targets.jl:37-144) and the ensemble closures of
 * build_ensemble_from_systems (src/Optimization/Ensemble/SystemBuilder.jl:49-87). */
typedef struct {
  int dim;                  /* d (closed) or N^2 (Lindblad); 1 <= dim <= 64, or
                               up to 128 for closed-dynamics gate problems under
                               PR_GRAD_REFERENCE (PR_ERR_UNSUPPORTED otherwise) */
  int n_ctrl;               /* number of control channels, >= 1 */
  int n_t;                  /* number of time slices, >= 1 */
  double dt;                /* uniform slice length (used when dt_slices == NULL) */
  const double* dt_slices;  /* [n_t] per-slice lengths (MR pulse_dt) or NULL */
  int dynamics;             /* PR_CLOSED or PR_LINDBLAD */
  int objective;            /* PR_GATE, PR_PAIRS_SQUARE or PR_PAIRS_REAL */
  int n_members;            /* ensemble members, >= 1 */
  const double* H0;         /* drift A0: [dim*dim] complex, or [n_members][dim*dim] */
  int H0_per_member;        /* 1: H0 holds one drift per member */
  const double* Hc;         /* controls A_j: [n_ctrl][dim*dim] complex */
  const double* ctrl_scale; /* [n_members][n_ctrl] scale s_mj, or NULL (all 1) */
  const double* weight;     /* [n_members] w_m, or NULL: 1/n_members (gate),
                               1/(n_members*n_pairs) (pairs) */
  const double* U_target;   /* PR_GATE: [dim*dim] complex */
  int n_pairs;              /* pair objectives: number of (init, target) pairs */
  const double* psi_init;   /* [n_pairs][dim] complex */
  const double* psi_targ;   /* [n_pairs][dim] complex (not read by PR_TRACKING) */
  /* PR_TRACKING checkpoints (TrackingPoint, GRAPETracking.jl:89-93); unused otherwise.
   * Checkpoint c compares rho after slice track_step[c] (1-based, 1..n_t) with
   * the density matrix track_target[c] at weight track_weight[c]; several
   * checkpoints may share a step. psi_init holds the initial pure states. */
  int n_track;                /* number of checkpoints, >= 1 for PR_TRACKING */
  const int* track_step;      /* [n_track] */
  const double* track_weight; /* [n_track] */
  const double* track_target; /* [n_track][dim*dim] complex */
} pr_problem;

#define PR_SEGMENT_AUTOTUNE (-1)

typedef struct {
  int precision;   /* PR_FP64 or PR_MIXED */
  int grad_mode;   /* PR_GRAD_REFERENCE or PR_GRAD_EXACT */
  int segment;     /* slices per time segment of the parallel chain scan; 0 = auto (about
                      8192 equal work items per dimension group); PR_SEGMENT_AUTOTUNE =
                      measure 0.5, 1, 2 and 4 times the auto length, each with the
                      single-wavefront kernels' trace operand and split assembly tables
                      in registers and read from memory, for every dimension group at
                      plan creation and keep the fastest (no reference counterpart: the
                      reference runs slices one after another) */
  int store_props; /* 1: keep U_k in HBM between the passes; 0: recompute U_k in the
                      gradient walk for FP64 problems (dim > 64 and Lindblad: from
                      segment checkpoints every 32 slices at most; Lindblad problems
                      also keep their forward states only for the work items of one
                      launch window, 8192 / ceil(dim/16)^2 of them; PR_MIXED and
                      PR_SPLIT16 problems rebuild their reduced-precision images,
                      tracking problems their U_k); -1: recompute
                      only when all images (and Lindblad forward states) would take
                      more than half of the free device memory */
  int pack;        /* 1: problems of dim <= 8 and objective PR_GATE, PR_PAIRS_SQUARE,
                      PR_PAIRS_REAL, PR_PAIRS_MODULUS, PR_STATE or PR_TRACKING (closed
                      dynamics), or PR_PAIRS_SQUARE or PR_PAIRS_REAL (Lindblad dynamics),
                      are packed
                      as diagonal blocks of problems of dim <= 16 (up to 16 rows and 16
                      pairs), so that one 16 x 16 work item evaluates several of them:
                      one-member problems without per-slice dt, of one dynamics,
                      objective and dt, together; the members of an ensemble, several to
                      a wider member (PR_TRACKING ensembles too, when every member has the
                      same weight: a divisor of the member count to a wider member).
                      F and G are each problem's own, with its weights, equal to an
                      unpacked plan's to rounding; the diagnostics and exp counts are the
                      packed problem's. 0: no packing. No
                      reference counterpart: the reference evaluates problems one by one.
                      A NULL opt means precision PR_FP64, grad_mode PR_GRAD_REFERENCE,
                      segment 0, store_props -1, pack 1. */
} pr_options;

/* Per-problem diagnostics filled by pr_eval when diag != NULL (a packed
 * problem, pr_options.pack, reports its pack's):
 * diag[4*p + 0] = max over members of ||U_total^dag U_total - I||_max (closed)
 *                 or 0 (Lindblad: the channel is not unitary)
 * diag[4*p + 1] = largest scaling exponent used by exp over the problem
 * diag[4*p + 2] = largest Taylor degree used by exp (6 or 9)
 * diag[4*p + 3] = number of time segments of the chain scan */
#define PR_DIAG_PER_PROBLEM 4

int pr_abi_version(void);
/* Number of visible devices; 0 when none. Never fails. */
int pr_device_count(void);
/* Select and initialise a device (gfx950 required). */
int pr_init(int device);
/* Copies the last error message of the calling thread (NUL-terminated). */
int pr_last_error(char* buf, size_t n);

/* Build a plan for n_probs problems. Validates every problem (Hermiticity of
 * H0/Hc to 1e-10 for closed dynamics as QuantumSystem does,
 * src/Types/QuantumSystem.jl:60-66; unitarity of U_target to 1e-8 as
 * unitary_target does, src/Types/Targets.jl:138-142), copies the operators to
 * the device and allocates all scratch. The host arrays may be freed after. */
int pr_plan_create(const pr_problem* probs, int n_probs, const pr_options* opt, void** plan);

/* Host-in / host-out evaluation. controls[p] is [n_ctrl x n_t] of problem p;
 * F[p] receives its fidelity; G[p] (may be NULL to skip the gradient for all
 * problems) receives [n_ctrl x n_t]; diag may be NULL. */
int pr_eval(void* plan, const double* const* controls, double* F, double* const* G, double* diag);

/* Device-resident evaluation on `stream` (a hipStream_t, or NULL for the
 * plan's own stream): d_controls holds every problem's controls back to back
 * in problem order, d_F receives n_probs fidelities and d_G (or NULL) the
 * gradients back to back in the same order. Asynchronous: returns after
 * enqueueing; all pointers are device pointers. No counterpart in the
 * reference (its GPU paths copy host arrays on every call); it serves callers
 * that keep controls resident in HBM across optimizer iterations. */
int pr_eval_device(void* plan, const double* d_controls, double* d_F, double* d_G, void* stream);

/* pr_eval_device with no allocation and no graph capture of its own: the same
 * kernels and results, with every scratch buffer of the evaluation (segment
 * products, prefixes, adjoint seeds, U_k images, per-member fidelities and
 * gradients, reduction partials) in a static array of 64 MiB per GPU instead
 * of the plan's memory. It has no counterpart in the reference. It is for
 * callers that cannot allocate device memory while they run: code under HIP
 * graph capture (the launches go straight onto `stream`, with fill kernels
 * instead of memsets), tight serving loops, harnesses that forbid allocation.
 * Limits: a plan whose scratch exceeds 64 MiB (see pr_options.store_props = 0
 * to drop the images) or that holds problems with dim > 64 is rejected with
 * PR_ERR_UNSUPPORTED. Every call on a GPU uses that GPU's one array, so calls
 * that overlap on one GPU (two streams, two threads) race: serialise them. The
 * array keeps the last call's contents; what must start at zero is cleared on
 * every call. pr_init must have been called for the device. */
int pr_eval_device_static(void* plan, const double* d_controls, double* d_F, double* d_G, void* stream);

/* The plan's own device buffers used by pr_eval (controls and gradient:
 * pr_plan_n_params doubles each; F: n_probs doubles). */
int pr_plan_buffers(void* plan, double** d_controls, double** d_F, double** d_G);
/* Copy host controls (as for pr_eval) into the plan's device control buffer. */
int pr_plan_upload_controls(void* plan, const double* const* controls);
/* Wait for all work enqueued on the plan's own stream. */
int pr_plan_synchronize(void* plan);

/* Total number of control values of the plan (sum of n_ctrl*n_t), i.e. the
 * length of d_controls / d_G for pr_eval_device. */
long long pr_plan_n_params(void* plan);

/* Exp work of the last evaluation (no counterpart in the reference; it
 * serves work models of the kernels): per problem p, counts[4p] member-slices
 * whose exp took the degree-6 polynomial, counts[4p + 1] those that took
 * degree 9, counts[4p + 2] the squarings over all member-slices, counts[4p + 3]
 * the member-slices a reduced-precision plan redid in FP64. Waits for the
 * plan's stream. Tracking problems report zeros; the problems of a pack
 * (pr_options.pack) report the pack's counts at its first problem and zeros
 * at the others. */
int pr_plan_exp_counts(void* plan, long long* counts);

/* Kernel shapes the plan uses (no counterpart in the reference; it reports
 * the choices of the segment rule or of PR_SEGMENT_AUTOTUNE): per problem p,
 * shapes[4p] slices per time segment, shapes[4p + 1] segments per member,
 * shapes[4p + 2] 1 when the single-wavefront gradient walk holds the trace
 * operand in registers (0: read from LDS), shapes[4p + 3] the assembly passes
 * the split segment kernel holds in registers (0: general assembly from
 * memory, or not a split problem); a packed problem reports its pack's. */
int pr_plan_shapes(void* plan, int* shapes);

/* Per-kernel timing (no counterpart in the reference, which has no device
 * profiling): when enabled, every evaluation records a HIP event pair on its
 * stream around each kernel launch. pr_plan_kernel_times waits for the
 * recorded events, returns the accumulated milliseconds and launch counts per
 * kernel class since the previous call, and resets them. Classes:
 * 0 segment exps + chain products, 1 prefix combine + overlaps,
 * 2 gradient walk + contraction, 3 ensemble reductions. */
#define PR_KERNEL_CLASSES 4
int pr_plan_set_timing(void* plan, int enable);
int pr_plan_kernel_times(void* plan, double* ms, long long* launches);

int pr_plan_destroy(void* plan);

/* ---- the GRAPE outer loop on the device (grape_optimize, GRAPE.jl:469-643) ----
 * Gradient ascent with the reference's step adaptation (adapt_grape_step_size,
 * GRAPE.jl:309-334: x1.05 when F rose, x0.5 when it fell by more than 1e-14,
 * clamped), update (grape_step!, GRAPE.jl:261-268, u += step G) and
 * convergence test (check_grape_convergence, GRAPE.jl:371-402), with the
 * controls, gradient and best controls resident in HBM: per iteration one
 * evaluation graph, the gradient norm and the update run on the device and
 * only F and |G| come back. fid_plan (one problem) evaluates F; grad_plan (one
 * problem, same controls layout) evaluates G, or NULL when fid_plan's
 * gradient is the one wanted (gate targets). */
typedef struct {
  int max_iter;
  double convergence_tol;
  double step_size;
  int adapt_step_size;
  double min_step_size;
  double max_step_size;
  double gradient_norm_tol;
} pr_grape_config;

#define PR_STOP_MAX_ITER 0
#define PR_STOP_PERFECT 1  /* F >= 1 - 1e-12 */
#define PR_STOP_GRADIENT 2 /* |G| < gradient_norm_tol */
#define PR_STOP_FIDELITY 3 /* |F_n - F_{n-1}| < convergence_tol (n >= 3 entries) */

typedef struct {
  int n_iterations;
  int converged;
  int reason;              /* PR_STOP_* */
  int n_fidelity_history;  /* entries written to fidelity_history */
  int n_fidelity_evaluations;
  int n_gradient_evaluations;
  double fidelity;         /* F at the best controls (re-evaluated, as the reference does) */
  double step_size_final;
} pr_grape_result;

/* controls_init / controls_best: [n_ctrl x n_t] column-major host arrays.
 * fidelity_history needs max_iter + 2 entries, gnorm_history max_iter. */
int pr_grape_optimize(void* fid_plan, void* grad_plan, const double* controls_init, const pr_grape_config* cfg,
                      double* controls_best, double* fidelity_history, double* gnorm_history,
                      pr_grape_result* result);

/* Ensemble L-BFGS-GRAPE with the iteration on the device: replaces the loop of
 * grape_optimize_ensemble (src/Optimization/GRAPE.jl:1012-1230), with the
 * two-loop recursion and safeguards of :1062-1080 and the backtracking line
 * search of :1082-1104 (F only, step halved up to 20 times, accepted at
 * F_new >= F). The curvature pairs (S, Y), the controls, the gradient and the
 * best controls stay in device memory; the host reads F, |g|, max|d| and the
 * pair's curvature each iteration. plan: one problem (the ensemble, weights
 * 1/M for the mean). Stops: PR_STOP_GRADIENT (|g| < gradient_norm_tol),
 * PR_STOP_PERFECT (F >= 1 - 1e-8), PR_STOP_FIDELITY (|F_n - F_{n-1}| <
 * convergence_tol), PR_STOP_ZERO_DIRECTION (max|d| < 1e-14). */
typedef struct {
  int max_iter;
  double convergence_tol;
  double step_size;          /* initial trial step (times amplitude_limit when it is > 0) */
  double gradient_norm_tol;
  int lbfgs_memory;          /* 1..64 curvature pairs */
  double amplitude_limit;    /* > 0: controls clipped to [-limit, limit]; <= 0: no clipping */
} pr_lbfgs_config;

#define PR_STOP_ZERO_DIRECTION 4

/* controls_init / controls_best: [n_ctrl x n_t] column-major host arrays;
 * fidelity_history needs max_iter + 1 entries, gnorm_history max_iter.
 * result->step_size_final is the last accepted line-search step. */
int pr_lbfgs_optimize(void* plan, const double* controls_init, const pr_lbfgs_config* cfg, double* controls_best,
                      double* fidelity_history, double* gnorm_history, pr_grape_result* result);

/* The MR optimcon loops on the device: replaces _optimcon_lbfgs /
 * _optimcon_grape and their tracking forms (src/Application/MR/OptControl.jl:
 * 380-484, 488-561, 1040-1216) for one plan problem built from an MRControl
 * or LindbladMRControl (state pairs, Lindblad or tracking objective).
 *   PR_MR_LBFGS: the two-loop recursion of OptControl.jl:565-590 (no
 *     safeguards), the direction divided by max|d| (stop PR_STOP_ZERO_DIRECTION
 *     below 1e-14), the Armijo backtracking of :594-611 from
 *     a0 = 0.05 (u_bound - l_bound) (accepted at F_new >= F - 1e-4 |a|, up to
 *     20 halvings, then the last trial if it improves F, else no move; F-only
 *     evaluations, counted as the reference counts them), the pair
 *     s = w_new - w, y = G_new - G kept when y.s > 1e-14 |s|^2, the oldest
 *     dropped beyond lbfgs_memory.
 *   PR_MR_GRAPE: w <- clip(w + 0.05 (u_bound - l_bound) / max|G| G).
 * Every trial point is clipped to [l_bound, u_bound], the guess too. Stops
 * after an iteration: PR_STOP_PERFECT (F >= 1 - 1e-8), PR_STOP_GRADIENT
 * (|G| < grad_tol), and for L-BFGS PR_STOP_FIDELITY (|F_n - F_{n-1}| < 1e-9,
 * "fidelity stalled"). The waveform, gradient, curvature pairs and best
 * waveform stay in device memory; per evaluation the host reads F (and |G|,
 * max|d|, the pair's curvature). tile > 1: the plan holds packed members
 * (f = tile diagonal blocks per member), its controls are tile copies of the
 * waveform rows ([tile n_ctrl x n_t]) and the waveform's gradient is the sum
 * of the tile channel rows. */
#define PR_MR_LBFGS 0
#define PR_MR_GRAPE 1
typedef struct {
  int method;        /* PR_MR_LBFGS or PR_MR_GRAPE */
  int max_iter;
  double grad_tol;
  int lbfgs_memory;  /* 1..64 */
  double l_bound;
  double u_bound;
  int tile;          /* >= 1 */
} pr_mr_config;

/* controls_init / controls_best: the waveform, [n_ctrl x n_t] column-major
 * (n_ctrl = the plan's n_ctrl / tile). fidelity_history and gnorm_history need
 * max_iter + 1 entries each (entry 0: the guess). result->fidelity is the best
 * F of the history, result->n_iterations the history's length - 1. */
int pr_mr_optimcon(void* plan, const double* controls_init, const pr_mr_config* cfg, double* controls_best,
                   double* fidelity_history, double* gnorm_history, pr_grape_result* result);

#ifdef __cplusplus
}
#endif

#endif /* PULSAR_ROCM_H */
