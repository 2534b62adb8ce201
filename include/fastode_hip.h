/* C ABI of the MI355X (gfx950) batched Rosenbrock integrator.
 *
 * Every entry point takes caller-owned device pointers (FP64 unless noted)
 * and a hipStream_t passed as void*; the library never frees caller memory.
 * Return value: 0 on success, a negative code on a bad argument or a HIP
 * error (fode_last_error() gives the text).  Numerical failures are not
 * errors: they come back as per-group status codes, as the reference returns
 * status strings (fastode/solvers.py:330-367).
 *
 * Layouts.  A group's state Y is (C, n) row-major (cell-major: one cell's
 * species contiguous), rate constants k are (C, R), Jacobian values are
 * (C, nnz) in the mechanism's pattern order, a stored factorization is
 * (C, lu_len).  One device workgroup owns one cell of the LU.
 *
 * Handles are not thread-safe; calls on one stream are ordered.
 */
#ifndef FASTODE_HIP_H
#define FASTODE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, mapped to the reference's strings by the host code. */
enum {
  FODE_RUNNING = 0,
  FODE_CONVERGED = 1,        /* "converged" */
  FODE_TOO_MANY_STEPS = 2,   /* "too many steps" */
  FODE_STEP_TOO_SMALL = 3,   /* "step too small" */
  FODE_NAN_OR_INF = 4,       /* "nan or inf" */
  FODE_PROJECTION_FAILED = 5, /* "projection failed" */
  FODE_ACCEPTED_UNCONVERGED = 6 /* "accepted unconverged" (backward Euler) */
};

/* Next action of a group after a control update: a Rosenbrock attempt with
 * new F0 and J (NEW_STEP) or with the step's F0 and J (RETRY), a Newton pass
 * of backward Euler (NEWTON), or none (FINISH). */
enum { FODE_FINISH = 0, FODE_NEW_STEP = 1, FODE_RETRY = 2, FODE_NEWTON = 3 };

/* The step loop a controller runs (fode_ctrl_params.loop). */
enum {
  FODE_LOOP_ALGORITHM = 0,   /* Algorithm.solve (fastode/solvers.py:270-415) */
  FODE_LOOP_RESTARTING = 1,  /* RestartingLoop.solve (fastode/solvers.py:189-267) */
  FODE_LOOP_BACKWARD_EULER = 2 /* BackwardEuler.solve (fastode/solvers.py:418-491) */
};

/* Backward Euler's state after a Newton pass (fode_group_ctrl.accepted):
 * the iterate becomes the step's start (COMMIT) or is reset to it (RESTORE). */
enum { FODE_BE_KEEP = 0, FODE_BE_COMMIT = 1, FODE_BE_RESTORE = 2 };

/* Operation counters, 13 x int64 per group, in the order of
 * fastode/counts.py:9-23. */
enum {
  FODE_C_STEPS = 0, FODE_C_ACCEPTED, FODE_C_REJECTED, FODE_C_RHS, FODE_C_JAC, FODE_C_LU,
  FODE_C_SOLVES, FODE_C_NEWTON, FODE_C_FAILURES, FODE_C_UNCONVERGED, FODE_C_PROJ_PASSES,
  FODE_C_PROJ_TRIALS, FODE_C_CALLS, FODE_NCOUNTS
};

/* Per-group controller state on the device: the local variables of
 * Algorithm.solve (fastode/solvers.py:305-415) plus the carried Memory
 * (fastode/solvers.py:181-186) and the group's Counts.  RestartingLoop uses
 * the same fields (no Memory); backward Euler names four of them anew. */
typedef struct fode_group_ctrl {
  double t, H, H_ctrl, proposed, dt, h_min, h_max, err, alpha;
  double h_next, err_prev, h_prev; /* Memory */
  int32_t nsteps, status;
  union { int32_t last; int32_t iterations; };      /* backward Euler: Newton passes of the step */
  union { int32_t shortened; int32_t converged; };  /* backward Euler: the last pass converged */
  union { int32_t rej_last; int32_t n_success; };   /* backward Euler: successes in a row */
  union { int32_t rej_more; int32_t n_fail; };      /* backward Euler: failed steps of the call */
  int32_t action, accepted;        /* accepted: backward Euler's FODE_BE_* */
  int32_t fails, first_fail;       /* batched form: failed calls, first failing call */
  int64_t counts[FODE_NCOUNTS];
} fode_group_ctrl;

/* Solver options and tableau constants: Algorithm's options
 * (fastode/solvers.py:275-292), RestartingLoop's (:194-201), BackwardEuler's
 * (:423-428); loop selects the step loop (FODE_LOOP_*). */
typedef struct fode_ctrl_params {
  double b, safety, fac_min, fac_max, k_exp, gamma;
  int32_t h211b, carry, quantize, carry_memory, max_steps, stages, n_new_forcing, constrained;
  double a[16], c[16], m[4], e[4];
  int32_t new_forcing[4];
  int32_t loop;               /* FODE_LOOP_* */
  int32_t step_change_rows;   /* RestartingLoop: algebraic rows of E = Ynew - Y */
  int32_t max_iterations;     /* backward Euler: Newton passes per step */
  int32_t n_reductions;       /* backward Euler: entries of reductions (<= 8) */
  double reductions[8];       /* backward Euler: step cuts after failures */
  double small;               /* backward Euler: least convergence bound */
} fode_ctrl_params;

/* Mechanism tables (see the package's mechanism.py). */
typedef struct fode_mech_desc {
  int32_t n, R, P, nnz, ncon;
  const int32_t* react;      /* R*P, pad = n */
  const int32_t* rhs_ptr;    /* n+1 */
  const int32_t* rhs_rxn;
  const double* rhs_coef;
  const int32_t* jac_ptr;    /* nnz+1 */
  const int32_t* jac_src;    /* j*P + p */
  const double* jac_coef;
  const double* jac_const;   /* nnz */
  const int32_t* con_rows;   /* ncon */
  const double* con_coef;    /* ncon*n */
  const double* con_total;   /* ncon */
} fode_mech_desc;

/* Symbolic LU plan (see the package's symbolic.py for the table formats).
 * Tasks are 2 x uint32 each; counts are in tasks, pairs or words. */
typedef struct fode_plan_desc {
  int32_t n, c, m, tile, ns, nsch, nnz, slot_stride;
  int32_t n_f_tasks, f_nst, n_f_pairs;
  int32_t n_sf_tasks, sf_nst, n_sb_tasks, sb_nst, n_s_pairs, n_sch_pairs;
  const int32_t* sp_src; const int32_t* sp_diag; const int32_t* sp_row;     /* ns */
  const int16_t* core_src;   /* 256 x slot_stride, thread-major */
  const uint32_t* sch_mask;  /* 256 x 4: bit s set when core slot s has a Schur sum */
  const int32_t* sch_base;   /* 257: each thread's Schur pairs [sch_base[t], sch_base[t+1]), a multiple of 4 */
  const uint32_t* sch_pairs; /* n_sch_pairs: a << 16 | b, bit 31 on the pair closing a sum */
  const int32_t* core_row;   /* 16 * tile */
  const uint32_t* f_tasks; const int32_t* f_ptr; const uint32_t* f_pairs;
  const uint32_t* sf_tasks; const int32_t* sf_ptr;
  const uint32_t* sb_tasks; const int32_t* sb_ptr;
  const uint32_t* s_pairs;
  const int32_t* perm;       /* n: new -> old */
} fode_plan_desc;

const char* fode_last_error(void);

/* Device properties the host needs (LDS per workgroup, CUs). */
int fode_device_info(int device, int* lds_per_block, int* cus, int* warp_size);

/* Mechanism: replaces Mechanism's compiled tables (fastode/chem.py:19-79). */
int fode_mech_create(const fode_mech_desc* desc, void** out);
int fode_mech_destroy(void* mech);

/* Plan: replaces SparsePattern + SuperLU's per-call symbolic work
 * (fastode/linalg.py:35-81). lu_len receives doubles per cell of a factor. */
int fode_plan_create(const fode_plan_desc* desc, void** out, int64_t* lu_len);
int fode_plan_destroy(void* plan);

/* F = rhs(Y): Mechanism.rhs (fastode/chem.py:86-90) with linear constraint
 * rows. Y (C,n), k (C,R), F (C,n). */
int fode_rhs(void* mech, int C, const double* Y, const double* k, double* F, void* stream);

/* J = jac(Y): Mechanism.jac_values (fastode/chem.py:92-102). J (C,nnz). */
int fode_jac(void* mech, int C, const double* Y, const double* k, double* J, void* stream);

/* LU of alpha*diag(mass) - J (or diag(mass) - J restricted to algebraic rows
 * when alg_only), per cell: SparsePattern.matrix_values + SparseLU /
 * dense_matrix + DenseLU (fastode/linalg.py:15-32, 60-81). alpha is a device
 * scalar. J holds nnz values per cell in the plan's pattern order. */
int fode_factor(void* plan, int C, const double* alpha, const double* mass, const double* J,
                int alg_only, double* LU, void* stream);

/* X = LU.solve(B) (fastode/linalg.py:21, 80-81). B, X (C,n); may alias. */
int fode_solve(void* plan, int C, const double* LU, const double* B, double* X, void* stream);

/* Controller, one group: ctrl_init = the set-up of Algorithm.solve
 * (fastode/solvers.py:306-320) and, unless begin_later, the first loop-top
 * checks; ctrl_begin = those checks (after a projection). */
int fode_ctrl_init(const fode_ctrl_params* p, fode_group_ctrl* g, int ngroups, int begin_later,
                   void* stream);
int fode_ctrl_begin(const fode_ctrl_params* p, fode_group_ctrl* g, int ngroups, void* stream);

/* Stage arithmetic of `stages` (fastode/solvers.py:54-86), group of C cells.
 * K holds the stage vectors, each (C,n). */
int fode_stage_point(const fode_ctrl_params* p, int r, int C, int n, const double* Y,
                     const double* const* K, double* out, void* stream);
int fode_stage_forcing(const fode_ctrl_params* p, int r, int C, int n, const fode_group_ctrl* g,
                       const double* F, const double* const* K, const double* mass, int dae,
                       double* out, void* stream);
int fode_stage_final(const fode_ctrl_params* p, int C, int n, const double* Y,
                     const double* const* K, double* Ynew, double* E, void* stream);

/* error_norm (fastode/solvers.py:37-43) over the group, then the
 * accept/reject update of Algorithm.solve (fastode/solvers.py:362-411) on
 * the device; on acceptance Y <- Ynew, on finishing Y is clamped
 * (fastode/solvers.py:89-97, 412-415).  work: C doubles of device memory
 * (the cells' shares of the norm, one workgroup per cell), or NULL for one
 * workgroup over the whole group; both give the same bits. */
int fode_norm_control(const fode_ctrl_params* p, fode_group_ctrl* g, int C, int n, double* Y,
                      const double* Ynew, const double* E, const double* atol, double rtol,
                      const double* mass, double* work, void* stream);

/* Backward Euler's Newton pass around the caller's rhs/jac/factor/solve
 * (BackwardEuler.solve, fastode/solvers.py:448-461): out = F - (Yn1 - Yn)/H
 * with H from g; then fode_be_update: Yn1 = max(Yn1 + delta, 0), the
 * group's convergence test (BackwardEuler.converged, :433-438) and the step
 * decision of :462-486 (the iterate kept, committed to Yn or reset to it),
 * clamping Yn1 when the call ends (:491).  The set-up of a call is
 * fode_ctrl_init with p->loop = FODE_LOOP_BACKWARD_EULER.  work: 2*C
 * doubles of device memory (each cell's flags, one workgroup per cell), or
 * NULL for one workgroup over the whole group; both give the same bits. */
int fode_be_residual(int C, int n, const fode_group_ctrl* g, const double* F, const double* Yn1,
                     const double* Yn, double* out, void* stream);
int fode_be_update(const fode_ctrl_params* p, fode_group_ctrl* g, int C, int n, double* Yn1, double* Yn,
                   const double* delta, const double* atol, double rtol, double* work, void* stream);

/* The per-interval photolysis rewrite of run_scenario (fastode/ts1.py:205-213)
 * on device rate constants: k[c*ldk + cols[q]] = (base[c*ncols + q] * f[0])
 * * scale[q] for c < C, q < ncols; base holds each cell's photolysis
 * parameters, scale the reactions' scaling factors, f (a device scalar) the
 * interval's factor. */
int fode_scale_columns(int C, int ldk, double* k, int ncols, const int32_t* cols, const double* base,
                       const double* scale, const double* f, void* stream);

/* clamp(system, Y) alone (fastode/solvers.py:89-97). */
int fode_clamp(int C, int n, double* Y, const double* mass, int constrained, void* stream);

/* ---- project (fastode/solvers.py:100-178), plain path: the arrays of the
 * damped Newton on the algebraic rows (mass == 0).  The host keeps the
 * reference's decisions and reads one measure per decision. */

/* out = where(algebraic, G, 0): the residual (fastode/solvers.py:121-124). */
int fode_proj_mask(int C, int n, const double* G, const double* mass, double* out, void* stream);

/* out = Y, plus step * D on the algebraic rows: a backtracking candidate
 * (fastode/solvers.py:151-153). */
int fode_proj_candidate(int C, int n, const double* Y, const double* D, double step, const double* mass, double* out,
                        void* stream);

/* Over the algebraic rows of V, into out[2] (device): out[0] = 1 if an entry
 * is not finite (fastode/solvers.py:126-127), out[1] = max |V| when Z is null
 * (the zero-residual test, :131-132), else the weighted max of
 * `weighted(Z, V)` (:112-118) with per-species atol and scalar rtol. */
int fode_proj_measure(int C, int n, const double* V, const double* Z, const double* mass, const double* atol,
                      double rtol, double* out, void* stream);

/* ---- Batched persistent form: every group of a batch over one host
 * interval in one launch.  It replaces the reference's group loop around
 * Algorithm.solve (fastode/ts1.py:196-216) with identical per-group
 * semantics: each group (group_size consecutive cells; the last may be
 * shorter) keeps its own controller state and Memory in g[group] across
 * calls.  Rate constants of cell c are k[c*k_stride + j] * kmul[j]
 * (k_stride = 0 shares one row); work is a C*n scratch (2*C*n when
 * p->constrained: the projection of fastode/solvers.py:100-178 runs per group
 * at the start of each call); queue is one int32 that the call resets itself.  Groups are taken from an atomic queue by
 * persistent workgroups, so uneven groups balance.  The launch sequence is
 * stream-ordered and free of host synchronization and memsets, so it may be
 * captured in a hipGraph. */
typedef struct fode_batch_desc {
  int32_t C, group_size, ngroups, k_stride;
  double dt, rtol;
  double* Y;               /* (C,n) state, updated in place */
  const double* k;         /* (C or 1, R) */
  const double* kmul;      /* (R) per-call multiplier (photolysis factor) */
  const double* atol;      /* (n) */
  const double* mass;      /* (n) */
  fode_group_ctrl* g;      /* (ngroups) */
  double* work;            /* (C,n); (2C,n) for a constrained tableau (projection) */
  int32_t* queue;          /* 1 int32 */
  /* Optional record of the accepted steps (for the adjoint), or NULL:
   * hlog[group*hcap + i] = H of the group's i-th accepted step of this call,
   * nacc[group] = their number (hcap >= the attempt budget, so it never
   * overflows). */
  double* hlog;
  int32_t* nacc;
  int32_t hcap;
  int32_t options;         /* bit 0: keep the plan's tables in memory (no LDS staging) */
  /* Optional order in which the persistent workgroups take groups (a
   * permutation of 0..ngroups-1), or NULL for 0, 1, 2, ...: costliest first
   * shortens the tail of an interval.  Results do not depend on it. */
  const int32_t* order;
} fode_batch_desc;

/* Launch one interval. blocks_per_cu <= 0 picks the occupancy limit. */
int fode_batched_interval(void* plan, void* mech, const fode_ctrl_params* p, const fode_batch_desc* b,
                          int blocks_per_cu, void* stream);

/* fode_batched_interval for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  b->work and b->queue must be NULL: the
 * library's own per-device scratch (32 MiB, a static device array, held by
 * each GPU until the process ends) takes their place, holding the queue
 * counter and one slot of group_size * n doubles (two with a projected
 * constrained tableau) per persistent workgroup; the launch runs at most as
 * many workgroups as there are slots, and rejects a group_size * n with no
 * slot at all.  Every other argument, and the results, are those of
 * fode_batched_interval.  Calls on one GPU share the scratch: calls that
 * overlap on one GPU (other streams, other threads) race. */
int fode_batched_interval_noalloc(void* plan, void* mech, const fode_ctrl_params* p, const fode_batch_desc* b,
                                  int blocks_per_cu, void* stream);

/* Phase timers of the batched kernel: the cycles its workgroups spent per
 * phase of a cell-attempt (load, rates, Jacobian, assembly, sparse factor,
 * Schur update, core factor, block inverses, stage set-up, sparse solves,
 * core solves, combination, control), summed over every launch since the
 * last reset, into out[0..n); reset clears them.  Returns the number of
 * phases, or -1 in a library built without FODE_PHASE_TIMERS (the product
 * build: `build.py --phase-timers` makes libfastode_hip_timed.so).  No
 * counterpart in the reference: a measurement facility.  Synchronous. */
int fode_phase_timers(uint64_t* out, int n, int reset);

/* LDS bytes one workgroup of the batched kernel uses for this plan/mechanism. */
int fode_batched_lds_bytes(void* plan, void* mech, int stages, int64_t* bytes);

/* The same interval launch captured once in a hipGraph (queue reset +
 * persistent kernel) for replay; the arrays named in b are baked in, their
 * contents (Y, kmul, g, ...) may change between replays. */
int fode_batched_graph_create(void* plan, void* mech, const fode_ctrl_params* p,
                              const fode_batch_desc* b, int blocks_per_cu, void** graph);
int fode_graph_launch(void* graph, void* stream);
int fode_graph_destroy(void* graph);

/* ---- Discrete adjoint of one interval of the batched form.  No counterpart
 * in the reference (its solver has no derivative); it differentiates the
 * map that fode_batched_interval applies, with the accepted steps that call
 * recorded (fode_batch_desc.hlog/nacc) held fixed: quantized steps are
 * locally constant in the inputs.  Per cell (independent given the steps):
 * the interval's accepted steps are recomputed from the checkpoint Y (the
 * state at the interval's start) and their start states and stages kept in
 * ws; then, last step first, (alpha*M - J)^T is factored with plan_t (the
 * plan of the transposed pattern under the same ordering) and the stage
 * adjoints, Hessian terms and rate-constant terms are accumulated:
 *   lam  (C,n): in, dL/d(state after the interval's clamp); out, dL/dY;
 *   kbar (C,R): += dL/d(log k) of this interval.
 * ws holds blocks * ws_stride doubles, ws_stride >= max(nacc)*(stages+1)*n;
 * fode_adjoint_ws_blocks tells the number of blocks.  Unconstrained systems
 * only. */
typedef struct fode_adjoint_desc {
  int32_t C, group_size, ngroups, k_stride, hcap, pad_;
  int64_t ws_stride;
  const double* Y;         /* (C,n) checkpoint */
  const double* k;         /* (C or 1, R) */
  const double* kmul;      /* (R) */
  const double* mass;      /* (n) */
  const double* hlog;      /* (ngroups, hcap) */
  const int32_t* nacc;     /* (ngroups) */
  double* lam;             /* (C,n) */
  double* kbar;            /* (C,R) */
  double* ws;              /* (blocks, ws_stride) */
  int32_t* queue;          /* 1 int32 */
} fode_adjoint_desc;

int fode_adjoint_interval(void* plan, void* plan_t, void* mech, const fode_ctrl_params* p,
                          const fode_adjoint_desc* a, void* stream);
/* Workgroups fode_adjoint_interval launches (size of ws in ws_stride units). */
int fode_adjoint_ws_blocks(void* plan, void* plan_t, void* mech, int stages, int C, int* blocks);

#ifdef __cplusplus
}
#endif
#endif
