/*
 * mpx_kkt.h -- C ABI of the MI355X (gfx950) primal-dual iLQR KKT hot path.
 *
 * The reference reaches its solver through two Python/JAX entry points of the
 * un-vendored submodule primal_dual_ilqr (.gitmodules:1-3):
 *   optimizers.mpc(cost, dynamics, hessian_approx, limited_memory, reference,
 *                  parameter, W, x0, X_in, U_in, V_in) -> (X, U, V)
 *       called at mpx/utils/mpc_wrapper.py:74, :150-158, :244, :346-354,
 *       mpx/utils/mpc_wrapper_srbd.py:68, :124-132, mpx/examples/acrobat.py:101-113
 *   optimizers.model_evaluator_helper(cost, dynamics, x0, X, U) -> (g, c)
 *       called at mpx/utils/mpc_wrapper.py:422, :452
 * and keeps its warm start with update_and_extract (mpx/utils/mpc_wrapper.py:79-96,
 * :261-282; mpx/utils/mpc_wrapper_srbd.py:135-137).  The functions below replace
 * those calls for a whole batch of instances (the jax.vmap of
 * mpx/utils/mpc_wrapper_srbd.py:77 and mpx/examples/multi_env.py:49) in one
 * stream-ordered call.
 *
 * Conventions
 *  - All pointers are device pointers (HBM) owned by the caller; nothing is
 *    allocated by the library except inside mpxkkt_* calls documented so.
 *  - Calls are enqueued on `stream` (a hipStream_t passed as void*; NULL is the
 *    default stream) and never synchronise the host.
 *  - Every entry point returns an mpxkkt_status; on error the reason is kept
 *    per thread in mpxkkt_last_error().  NaNs in the data propagate and are
 *    reported per instance in mpxkkt_stats.status, never as an error code.
 *  - Iterates are dense, instance-major, row-major:
 *      X [batch][N+1][n], U [batch][N][m], V [batch][N+1][n],
 *      Y_eq [batch][N][p], Y_box [batch][N][2m] (upper bounds, then lower).
 *  - Problem data is given as strided views: element (b, t, ...) of a view is
 *    at ptr + b*batch_stride + t*stage_stride (in elements), the trailing
 *    matrix row-major and dense.  A zero stride shares data across
 *    instances or stages (time-invariant Q, R; per-instance A, B, W).
 */
#ifndef MPX_KKT_H
#define MPX_KKT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  MPXKKT_OK = 0,
  MPXKKT_ERR_ARG = 1,          /* invalid argument (dims, null pointer, dtype) */
  MPXKKT_ERR_UNSUPPORTED = 2,  /* shape exceeds the device limits of a kernel */
  MPXKKT_ERR_WORKSPACE = 3,    /* workspace too small */
  MPXKKT_ERR_DEVICE = 4        /* HIP runtime error (launch, attribute) */
} mpxkkt_status;

enum { MPXKKT_F32 = 0, MPXKKT_F64 = 1 };

/* dynamics families evaluated on the device */
enum {
  MPXKKT_MODEL_LTV = 0,   /* x' = A_t x + B_t u                                   */
  MPXKKT_MODEL_TANH = 1,  /* x' = x + h (A x + B u + kappa tanh(W x))            */
  MPXKKT_MODEL_CALLER = 2 /* the caller's own dynamics and cost, linearised by the
                             caller at the iterate (see mpxkkt_problem) and
                             evaluated by the caller at the line-search trials:
                             mpxkkt_sqp_direction + mpxkkt_sqp_accept             */
};

typedef struct {
  const void* ptr;
  int64_t batch_stride;
  int64_t stage_stride;
} mpxkkt_view;

/* One batch of same-shaped instances.  Cost of the device model families
 * (LTV, TANH): sum_t 1/2 |x_t - xref_t|_Q_t^2 + 1/2 |u_t - uref_t|_R_t^2 +
 * 1/2 |x_N - xref_N|_QN^2 (the Gauss-Newton (Q, R, M = 0) triple of
 * hessian_approx, mpx/utils/objectives.py:75), with optional rows
 * D_t x + E_t u + e_t = 0 and boxes ulo <= u <= uhi.
 *
 * MPXKKT_MODEL_CALLER: the caller's dynamics(x, u, t, parameter)
 * (mpx/utils/models.py:6) and cost(W, reference, x, u, t)
 * (mpx/utils/objectives.py:15) stay with the caller, who passes what the
 * reference solver takes from jax.jacobian / jax.grad and from
 * hessian_approx(W, reference, x, u, t) -> (Q, R, M) (objectives.py:75, :202)
 * at the current iterate (X, U), per stage t < N unless noted:
 *   A = f_x (n x n), B = f_u (n x m), fbar = f(x_t, u_t) (n),
 *   Q (n x n), R (m x m), M (n x m, NULL = 0): the stage Hessian
 *   [[Q, M], [M', R]] (used as given; D2 floors the m x m pivots),
 *   q = grad_x l_t (n, rows 0..N), r = grad_u l_t (m),
 *   g = l_t (1, rows 0..N), QN = terminal Hessian (n x n).
 * xref, uref, W, h, kappa are unused.  The equality rows and boxes stay the
 * library's (their augmented-Lagrangian and barrier terms are added on the
 * device).  With boxes, linearise at U moved inside the box the way the
 * library does on entry: a coordinate u >= hi becomes hi - 1e-3 (hi - lo),
 * u <= lo becomes lo + 1e-3 (hi - lo). */
typedef struct {
  int32_t dtype;        /* MPXKKT_F32 / MPXKKT_F64 for every array */
  int32_t model;        /* MPXKKT_MODEL_* */
  int32_t batch, n, m, horizon, n_eq;
  int32_t has_box;
  double h, kappa;      /* MPXKKT_MODEL_TANH constants */
  mpxkkt_view A, B, W;  /* LTV: A_t (n x n), B_t (n x m); TANH: A, B, W per instance */
  mpxkkt_view Q, R, QN, xref, uref;
  mpxkkt_view D, E, e;  /* p x n, p x m, p (NULL when n_eq == 0) */
  mpxkkt_view ulo, uhi; /* m (NULL when has_box == 0) */
  mpxkkt_view x0;       /* n */
  mpxkkt_view fbar, M, q, r, g;  /* MPXKKT_MODEL_CALLER only (else NULL) */
} mpxkkt_problem;

typedef struct {
  void* X;
  void* U;
  void* V;
  void* Y_eq;   /* may be NULL when n_eq == 0 */
  void* Y_box;  /* may be NULL when has_box == 0 */
} mpxkkt_iterate;

typedef struct {
  double rho0, rho_max;   /* merit penalty floor and cap                    */
  double rho_c;           /* augmented-Lagrangian penalty of equality rows  */
  double mu_b, tau;       /* box barrier and fraction-to-boundary factor   */
  double beta, c1;        /* backtracking factor and Armijo constant       */
  double armijo_slack;    /* relative noise floor of the Armijo test       */
  int32_t max_trials;     /* line-search trials                            */
  int32_t limited_memory; /* 1: evaluate one trial at a time               */
  int32_t segments;       /* time-parallel scan segments per instance (1 = sequential Riccati) */
  int32_t block_threads;  /* threads per block of the per-instance kernels (0 = auto) */
  int32_t trial_group;    /* trials evaluated together (0 = all); 1 if limited_memory */
  int32_t flags;          /* bit 0: FP32 products on VALU instead of the matrix cores;
                             bit 1: no register-resident wave LQR kernel (FP32,
                             n + 1 <= 48, m <= 16, n_eq <= 16, one segment);
                             bit 2: single-kernel line search (one block per
                             instance) instead of the stage-parallel one;
                             bit 3: linearisation parallel over stages (lower
                             latency for small batches; n, m, n_eq <= 64);
                             bit 4: the row-per-lane linearisation and
                             line-search trials and the S-based forward pass
                             and costates instead of the matrix-core and
                             adjoint forms (FP32, tanh model with
                             stage-invariant A, W, B, Q, R, n <= 48, m <= 16,
                             n_eq <= 16), and in the wave LQR kernel the
                             homogeneous offset n on the matrix cores (when
                             n % 4 == 0 it is otherwise added on the VALU) */
  double hess_floor;      /* pivot floor of the m x m Cholesky factorisations
                             (H_t = R + B'S B, ...): pivot j is raised to
                             hess_floor * |H_jj| when it falls below; keeps
                             the step defined for indefinite stage Hessians */
} mpxkkt_params;

/* Per-instance outputs of one SQP iteration (each may be NULL). */
typedef struct {
  void* alpha;     /* accepted step            [batch] dtype */
  void* rho;       /* merit penalty            [batch] dtype */
  void* merit0;    /* phi(0)                   [batch] dtype */
  void* dmerit;    /* phi'(0)                  [batch] dtype */
  int32_t* status; /* 0 ok, 1 non-finite, 2 line search exhausted [batch] */
  int32_t* trials; /* trials examined: backtracking in groups of trial_group [batch] */
} mpxkkt_stats;

/* Defaults of the oracle's decisions D3-D6. */
mpxkkt_status mpxkkt_default_params(mpxkkt_params* out);

/* Bytes of device workspace mpxkkt_sqp_iterate / mpxkkt_kkt_residual need. */
mpxkkt_status mpxkkt_workspace_size(const mpxkkt_problem* prob, const mpxkkt_params* prm,
                                    size_t* bytes);

/* One primal-dual SQP iteration for every instance, in place on `it`:
 * linearise + augment (K1), LQR by sequential Riccati or the chunked
 * associative scan (K2/K3), multiplier step, merit line search and update (K4).
 * Replaces optimizers.mpc (mpx/utils/mpc_wrapper.py:74, :150-158). */
mpxkkt_status mpxkkt_sqp_iterate(void* stream, const mpxkkt_problem* prob,
                                 const mpxkkt_params* prm, mpxkkt_iterate* it,
                                 mpxkkt_stats* stats, void* workspace, size_t workspace_bytes);

/* mpxkkt_sqp_iterate without a workspace argument and without any allocation:
 * the workspace is a static device array of the library (one copy of
 * MPXKKT_NOALLOC_WORKSPACE_BYTES per device, held until the process ends).
 * No counterpart in the reference: it is for callers that cannot allocate --
 * code under HIP graph capture, serving loops and harnesses that forbid
 * hipMalloc.  Problems whose mpxkkt_workspace_size exceeds the array are
 * refused (MPXKKT_ERR_WORKSPACE).  Every call on one device shares that
 * device's array and no kernel reads workspace it has not written in the same
 * call, so nothing needs clearing between calls; calls that overlap on one
 * device (other streams, other threads) race and must be serialised by the
 * caller.  No phase events are recorded. */
#define MPXKKT_NOALLOC_WORKSPACE_BYTES (64u << 20)
mpxkkt_status mpxkkt_sqp_iterate_noalloc(void* stream, const mpxkkt_problem* prob,
                                         const mpxkkt_params* prm, mpxkkt_iterate* it,
                                         mpxkkt_stats* stats);

/* The caller-dynamics SQP iteration (MPXKKT_MODEL_CALLER) in two calls, the
 * merit line search split where the reference solver evaluates the caller's
 * functions at the trial points:
 *  1. mpxkkt_sqp_direction: augmentation of the caller's linearisation (K1),
 *     LQR (K2/K3), multiplier steps, merit penalty rho, phi(0), phi'(0) and the
 *     fraction-to-boundary cap a_max.  Writes the step dX [batch][N+1][n],
 *     dU [batch][N][m] and a_max [batch] (dtype) into `dir`.
 *  2. the caller evaluates, for trials j < n_trials at a_j = a_max beta^j,
 *     x = X + a_j dX, u = U + a_j dU:
 *        trial_g [batch][n_trials][N+1] = l_t(x_t, u_t) (t = N: terminal),
 *        trial_f [batch][n_trials][N][n] = f(x_t, u_t, t)
 *     (all candidates at once: the vmap over candidates of the reference).
 *  3. mpxkkt_sqp_accept: the trials' merits, the first a_j that passes the
 *     Armijo test (or the smallest, status 2), the multiplier and iterate
 *     update in place on `it`, stats.
 * Between the calls the workspace holds the step: no other call may use it. */
typedef struct {
  void* dX;         /* [batch][N+1][n] dtype */
  void* dU;         /* [batch][N][m]   dtype */
  void* alpha_max;  /* [batch]         dtype */
} mpxkkt_direction;

mpxkkt_status mpxkkt_sqp_direction(void* stream, const mpxkkt_problem* prob, const mpxkkt_params* prm,
                                   mpxkkt_iterate* it, mpxkkt_direction* dir, void* workspace,
                                   size_t workspace_bytes);

mpxkkt_status mpxkkt_sqp_accept(void* stream, const mpxkkt_problem* prob, const mpxkkt_params* prm,
                                mpxkkt_iterate* it, int32_t n_trials, const void* trial_g,
                                const void* trial_f, mpxkkt_stats* stats, void* workspace,
                                size_t workspace_bytes);

/* The single-rigid-body quadruped of the reference's SRBD controller,
 * evaluated on the device: dynamics quadruped_srbd_dynamics
 * (mpx/utils/models.py:6-37), cost quadruped_srbd_obj and Hessian triple
 * quadruped_srbd_hessian_gn (mpx/utils/objectives.py:7-75), bound as at
 * mpx/utils/mpc_wrapper_srbd.py:58-68.  State x = [p, quat (w first), dp,
 * omega] (n = 13), control = ground reaction forces of 4 feet (m = 12).
 * Per instance: W (24 x 24 residual weight, stage stride 0), reference
 * [N+1][17] = [p_ref, quat_ref, dp_ref, omega_ref, contact (4)], parameter
 * [N+1][16] = [contact (4), p_legs (12)]. */
typedef struct {
  double mass, dt;
  double inertia[9], inertia_inv[9];  /* row-major */
  mpxkkt_view W, reference, parameter;
} mpxkkt_srbd;

/* One SQP iteration of the SRBD problem with no host involvement: the
 * linearisation (analytic Jacobians, gradient, Gauss-Newton Hessian triple)
 * into the caller views of `prob` (MPXKKT_MODEL_CALLER, n = 13, m = 12; its
 * A, B, fbar, Q, R, QN, q, r, g -- and M when non-NULL -- must point at
 * writable device buffers, overwritten here), then mpxkkt_sqp_direction's
 * work, the model's cost and dynamics at the max_trials trial points, and
 * mpxkkt_sqp_accept's work.  Replaces the jit(vmap(optimizers.mpc)) call of
 * the SRBD controller (mpx/utils/mpc_wrapper_srbd.py:68, :77). */
mpxkkt_status mpxkkt_srbd_iterate(void* stream, const mpxkkt_problem* prob, const mpxkkt_srbd* model,
                                  const mpxkkt_params* prm, mpxkkt_iterate* it, mpxkkt_stats* stats,
                                  void* workspace, size_t workspace_bytes);

/* Gait of the SRBD controller's reference generator (mpx/config/config_srbd.py). */
typedef struct {
  double dt, mass;           /* MPC step (s) and body mass (kg)                      */
  double duty_factor, step_freq, step_height, clearance_speed;
  double foot0[12];          /* nominal foot positions in the base frame             */
  double advance_dt;         /* > 0: first advance t_timer by timer_run over this
                                step (1 / mpc_frequency), as the controller does
                                each tick (mpx/utils/mpc_wrapper_srbd.py:106)         */
} mpxkkt_srbd_gait;

/* The SRBD controller's gait timer and reference generator on the device,
 * reference_generator_srbd without the terrain estimator
 * (mpx/utils/mpc_utils.py:7-14, :146-268): per instance, from the gait
 * phases t_timer [4] (in/out), the state x [13], the feet [12], the command
 * input [7] = [v, omega, base height], the measured contacts [4] and the
 * liftoff positions [12] (in/out), the reference [N+1][17] and parameter
 * [N+1][16] mpxkkt_srbd_iterate reads, and the feet's reference velocities
 * foot_ref_dot [N+1][12] (may be NULL).  All arrays dense, instance-major,
 * of `dtype`. */
mpxkkt_status mpxkkt_srbd_reference(void* stream, int32_t dtype, int32_t batch, int32_t horizon,
                                    const mpxkkt_srbd_gait* gait, void* t_timer, const void* x,
                                    const void* foot, const void* input, const void* contact,
                                    void* liftoff, void* reference, void* parameter, void* foot_ref_dot);

/* Options of mpxkkt_solve: the iteration protocol of runOffline
 * (mpx/utils/mpc_wrapper.py:425-467). */
typedef struct {
  int32_t max_iter;    /* at most max_iter + 1 iterations (runOffline stops when i > max_iter) */
  int32_t check_every; /* 0: enqueue all max_iter + 1 iterations with no host synchronisation
                          (stream-ordered, capturable in a HIP graph; converged instances skip
                          the kernels of later iterations); k > 0: after every k iterations wait
                          for the stream and stop enqueueing once no instance is running */
  double dcost_tol;    /* stop when last - l2 < dcost_tol (1e-3) ...                      */
  double cc_tol;       /* ... and sum |c|^2 < cc_tol (1e-5), l2 = sum_t g_t^2              */
} mpxkkt_solve_opts;

mpxkkt_status mpxkkt_default_solve_opts(mpxkkt_solve_opts* out);

/* SQP iterations until runOffline's stop rule holds for every instance
 * (mpx/utils/mpc_wrapper.py:425-467, the loop of optimizers.mpc calls and
 * optimizers.model_evaluator_helper at :441-466): after iteration i (from 1)
 * (g, c) = model evaluation at the new iterate, l2 = sum g^2, and an
 * instance stops when i > max_iter or (last - l2 < dcost_tol and
 * sum c^2 < cc_tol), with last = 1e10 before the first iteration and l2 after
 * each later one.  A stopped instance's iterate and stats are left as they
 * are.  Outputs (device, [batch]): iters (int32, iterations run), cost (l2 of
 * the last iteration, dtype, may be NULL), cc (sum c^2, dtype, may be NULL).
 * `launched` (host, may be NULL) receives the iterations enqueued.  Same
 * workspace as mpxkkt_sqp_iterate. */
mpxkkt_status mpxkkt_solve(void* stream, const mpxkkt_problem* prob, const mpxkkt_params* prm,
                           const mpxkkt_solve_opts* opts, mpxkkt_iterate* it, mpxkkt_stats* stats,
                           int32_t* iters, void* cost, void* cc, void* workspace, size_t workspace_bytes,
                           int32_t* launched);

/* KKT residual 2-norm per instance at `it` (out: [batch], dtype).  The
 * stop test of runOffline, mpx/utils/mpc_wrapper.py:452-466, in KKT form. */
mpxkkt_status mpxkkt_kkt_residual(void* stream, const mpxkkt_problem* prob,
                                  const mpxkkt_params* prm, const mpxkkt_iterate* it,
                                  void* out, void* workspace, size_t workspace_bytes);

/* (g, c) of optimizers.model_evaluator_helper (mpx/utils/mpc_wrapper.py:422):
 * g [batch][N+1] per-stage cost, c [batch][N+1][n] = [x0 - X_0, f(x_t,u_t) - x_{t+1}]. */
mpxkkt_status mpxkkt_model_eval(void* stream, const mpxkkt_problem* prob, const void* X,
                                const void* U, void* g, void* c);

/* Warm-start shift with per-instance NaN fallback, update_and_extract of
 * mpx/utils/mpc_wrapper.py:79-96: if U[b,0,0] is NaN the instance is reset to
 * X_t = x0, U_t = u_reset (or 0), V = Y = 0 and reset[b] = 1; else every array
 * drops its first `shift` rows and repeats its last row. */
mpxkkt_status mpxkkt_shift_warmstart(void* stream, const mpxkkt_problem* prob, int32_t shift,
                                     mpxkkt_iterate* it, const void* u_reset,
                                     int32_t* reset);

/* Phase timing.  While enabled, every mpxkkt_sqp_iterate records HIP events on
 * its stream at the phase boundaries: linearise | LQR recursion (backward
 * Riccati or scan and the forward state chain) | costates and step
 * assembly | line search.  mpxkkt_profile_collect waits for the recorded
 * events, returns the summed milliseconds per phase and the number of calls,
 * and clears the record. */
#define MPXKKT_PHASES 4
mpxkkt_status mpxkkt_profile_enable(int32_t on);
mpxkkt_status mpxkkt_profile_collect(double ms_per_phase[MPXKKT_PHASES], int32_t* calls);

const char* mpxkkt_last_error(void);
int32_t mpxkkt_version(void);

#ifdef __cplusplus
}
#endif
#endif /* MPX_KKT_H */
