// Step control of Algorithm.solve on the device and the group error norm.
#pragma once
#include "common.hpp"

namespace fode {

// The step control of Algorithm.solve (fastode/solvers.py:297-415) as a
// state machine over fode_group_ctrl; run by one thread per group.

__device__ inline double quantize(double h, double interval) {
  if (h >= interval) return interval;
  double nn = ceil(interval / h - 1.0e-9);
  return interval / fmax(nn, 1.0);
}

__device__ inline void ctrl_finish(fode_group_ctrl& g) {
  g.h_next = g.proposed;
  if (g.status == FODE_RUNNING) g.status = FODE_CONVERGED;
  g.action = FODE_FINISH;
}

// Top of the rejection loop (solvers.py:350-360): checks, then an attempt.
__device__ inline void ctrl_inner(const fode_ctrl_params& p, fode_group_ctrl& g, int action) {
  if (g.nsteps >= p.max_steps) {
    g.status = FODE_TOO_MANY_STEPS;
    ctrl_finish(g);
    return;
  }
  if (g.H < g.h_min || g.t + 0.1 * g.H == g.t) {
    g.status = FODE_STEP_TOO_SMALL;
    ctrl_finish(g);
    return;
  }
  g.alpha = 1.0 / (g.H * p.gamma);
  g.counts[FODE_C_LU] += 1;
  g.counts[FODE_C_SOLVES] += p.stages;
  g.counts[FODE_C_RHS] += p.n_new_forcing;
  g.action = action;
}

// ---- RestartingLoop (fastode/solvers.py:189-267): no memory, no landing
// rule, the elementary factor with fixed constants, budget checked as
// nsteps > max_steps at the loop top only, and no checks before a retry.
__device__ inline void restart_attempt(const fode_ctrl_params& p, fode_group_ctrl& g, int action) {
  g.alpha = 1.0 / (g.H * p.gamma);
  g.counts[FODE_C_LU] += 1;
  g.counts[FODE_C_SOLVES] += p.stages;
  g.counts[FODE_C_RHS] += p.n_new_forcing;
  g.action = action;
}

__device__ inline void restart_finish(fode_group_ctrl& g) {
  if (g.status == FODE_RUNNING) g.status = FODE_CONVERGED;
  g.action = FODE_FINISH;
}

// Loop top (solvers.py:226-236).
__device__ inline void restart_outer(const fode_ctrl_params& p, fode_group_ctrl& g) {
  if (!((g.t - g.dt + EPS) <= 0.0) || g.status != FODE_RUNNING) {
    restart_finish(g);
    return;
  }
  if (g.nsteps > p.max_steps) {
    g.status = FODE_TOO_MANY_STEPS;
    restart_finish(g);
    return;
  }
  if ((g.t + 0.1 * g.H == g.t) || g.H <= fmin(EPS, g.h_min)) {
    g.status = FODE_STEP_TOO_SMALL;
    restart_finish(g);
    return;
  }
  g.H = fmin(g.H, fabs(g.dt - g.t));
  g.counts[FODE_C_RHS] += 1;
  g.counts[FODE_C_JAC] += 1;
  restart_attempt(p, g, FODE_NEW_STEP);
}

// Set-up (solvers.py:208-220).
__device__ inline void restart_init(fode_group_ctrl& g) {
  g.counts[FODE_C_CALLS] += 1;
  g.h_min = 1.0e-15 * g.dt;
  g.h_max = g.dt;
  g.H = fmin(fmax(g.h_min, 1.0e-6 * g.dt), g.h_max);
  g.proposed = g.H;
  g.t = 0.0;
  g.rej_last = g.rej_more = 0;
  g.nsteps = 0;
  g.status = FODE_RUNNING;
  g.accepted = 0;
  g.err = 0.0;
}

// After an attempt (solvers.py:244-262).
__device__ inline void restart_after(const fode_ctrl_params& p, fode_group_ctrl& g, double err) {
  g.err = err;
  g.accepted = 0;
  g.nsteps += 1;
  g.counts[FODE_C_STEPS] += 1;
  if (isnan(err) || isinf(err)) {
    g.status = FODE_NAN_OR_INF;
    restart_finish(g);
    return;
  }
  const double fac = fmin(6.0, fmax(0.2, 0.9 / pow(err, 1.0 / p.k_exp)));
  double Hn = g.H * fac;
  if (err < 1.0 || g.H < g.h_min) {
    g.counts[FODE_C_ACCEPTED] += 1;
    g.t = g.t + g.H;
    g.accepted = 1;
    Hn = fmax(g.h_min, fmin(Hn, g.h_max));
    if (g.rej_last) Hn = fmin(Hn, g.H);
    g.rej_last = g.rej_more = 0;
    g.H = Hn;
    restart_outer(p, g);
  } else {
    if (g.rej_more) Hn = g.H * 0.1;
    g.rej_more = g.rej_last;
    g.rej_last = 1;
    g.H = Hn;
    g.counts[FODE_C_REJECTED] += 1;
    restart_attempt(p, g, FODE_RETRY);
  }
}

// ---- BackwardEuler (fastode/solvers.py:418-491).  One Newton pass is a
// FODE_NEWTON action over the whole group; after it the kernel reports
// whether every cell's correction met the bound (conv) and whether the new
// iterate is finite, and be_after decides.

// Start of a step, or the end of the call (solvers.py:444-447).
__device__ inline void be_step(fode_group_ctrl& g) {
  if (!(g.t < g.dt)) {
    g.action = FODE_FINISH;
    return;
  }
  g.iterations = 0;
  g.converged = 0;
  g.action = FODE_NEWTON;
}

// A Newton pass is about to run: its counts and the stage-matrix scale.
__device__ inline void be_pass(fode_group_ctrl& g) {
  g.alpha = 1.0 / g.H;
  g.counts[FODE_C_NEWTON] += 1;
  g.counts[FODE_C_RHS] += 1;
  g.counts[FODE_C_JAC] += 1;
  g.counts[FODE_C_LU] += 1;
  g.counts[FODE_C_SOLVES] += 1;
}

__device__ inline void be_init(fode_group_ctrl& g) {
  g.counts[FODE_C_CALLS] += 1;
  g.H = g.dt;
  g.t = 0.0;
  g.n_success = g.n_fail = 0;
  g.nsteps = 0;
  g.status = FODE_CONVERGED;
  g.accepted = FODE_BE_KEEP;
  be_step(g);
  if (g.action == FODE_NEWTON) be_pass(g);
}

// After a Newton pass (solvers.py:448-486).  conv: every correction of the
// group within its bound and finite; finite: the new iterate is finite.
__device__ inline void be_after(const fode_ctrl_params& p, fode_group_ctrl& g, bool conv, bool finite) {
  g.accepted = FODE_BE_KEEP;
  g.iterations += 1;
  const bool ok = g.iterations > 1 && conv;
  g.converged = ok;
  if (!ok && g.iterations < p.max_iterations) {
    be_pass(g);
    return;
  }
  g.nsteps += 1;
  g.counts[FODE_C_STEPS] += 1;
  if (!ok) {
    g.counts[FODE_C_REJECTED] += 1;
    g.counts[FODE_C_FAILURES] += 1;
    g.n_success = 0;
    if (g.n_fail >= p.n_reductions) {
      g.t = g.t + g.H;
      g.counts[FODE_C_UNCONVERGED] += 1;
      g.status = finite ? FODE_ACCEPTED_UNCONVERGED : FODE_NAN_OR_INF;
      g.action = FODE_FINISH;
      return;
    }
    g.accepted = FODE_BE_RESTORE;
    g.H = g.H * p.reductions[g.n_fail];
    g.n_fail += 1;
  } else {
    g.counts[FODE_C_ACCEPTED] += 1;
    g.t = g.t + g.H;
    g.accepted = FODE_BE_COMMIT;
    g.n_success += 1;
    if (g.n_success >= 2) {
      g.n_success = 0;
      g.H = g.H * 2.0;
    }
  }
  g.H = fmin(g.H, g.dt - g.t);
  be_step(g);
  if (g.action == FODE_NEWTON) be_pass(g);
}

// np.maximum(x, 0.0): NaN stays NaN (backward Euler's iterate, solvers.py:458).
__device__ inline double np_max0(double x) { return (x >= 0.0 || isnan(x)) ? x : 0.0; }

// Top of the step loop (solvers.py:330-348).  Backward Euler opens its
// steps itself (be_init, be_after): nothing to do for it here.
__device__ inline void ctrl_outer(const fode_ctrl_params& p, fode_group_ctrl& g) {
  if (p.loop == FODE_LOOP_BACKWARD_EULER) return;
  if (p.loop == FODE_LOOP_RESTARTING) {
    restart_outer(p, g);
    return;
  }
  if (!((g.t - g.dt + EPS) <= 0.0) || g.status != FODE_RUNNING) {
    ctrl_finish(g);
    return;
  }
  if (g.nsteps >= p.max_steps) {
    g.status = FODE_TOO_MANY_STEPS;
    ctrl_finish(g);
    return;
  }
  if ((g.t + 0.1 * g.H == g.t) || g.H < g.h_min) {
    g.status = FODE_STEP_TOO_SMALL;
    ctrl_finish(g);
    return;
  }
  const double rem = g.dt - g.t;
  g.last = rem <= g.H * (1.0 + 1.0e-8);
  g.shortened = g.last && rem < g.H * (1.0 - 1.0e-8);
  g.H_ctrl = g.H;
  if (g.last) g.H = rem;
  g.counts[FODE_C_RHS] += 1;
  g.counts[FODE_C_JAC] += 1;
  ctrl_inner(p, g, FODE_NEW_STEP);
}

// Set-up of a call (solvers.py:306-329); RestartingLoop's and backward
// Euler's set-ups by p.loop (backward Euler's also opens its first step).
__device__ inline void ctrl_init(const fode_ctrl_params& p, fode_group_ctrl& g) {
  if (p.loop == FODE_LOOP_RESTARTING) {
    restart_init(g);
    return;
  }
  if (p.loop == FODE_LOOP_BACKWARD_EULER) {
    be_init(g);
    return;
  }
  if (!(p.carry && p.carry_memory)) { g.err_prev = 0.0; g.h_prev = 0.0; }
  g.counts[FODE_C_CALLS] += 1;
  g.h_min = 1.0e-15 * g.dt;
  g.h_max = g.dt;
  double H = (p.carry && g.h_next > 0.0) ? g.h_next : 1.0e-6 * g.dt;
  H = fmin(fmax(g.h_min, H), g.h_max);
  if (p.quantize) H = quantize(H, g.dt);
  g.H = H;
  g.proposed = H;
  g.t = 0.0;
  g.rej_last = g.rej_more = 0;
  g.nsteps = 0;
  g.status = FODE_RUNNING;
  g.accepted = 0;
  g.err = 0.0;
}

__device__ inline double ctrl_factor(const fode_ctrl_params& p, const fode_group_ctrl& g, double err) {
  double fac = p.safety / pow(err, 1.0 / p.k_exp);
  if (p.h211b && g.err_prev > 0.0 && g.h_prev > 0.0)
    fac = p.safety * pow(err * g.err_prev, -1.0 / (p.b * p.k_exp)) * pow(g.H / g.h_prev, -1.0 / p.b);
  return fmin(p.fac_max, fmax(p.fac_min, fac));
}

// After an attempt with error err (solvers.py:362-411). Sets g.accepted.
__device__ inline void ctrl_after(const fode_ctrl_params& p, fode_group_ctrl& g, double err) {
  if (p.loop == FODE_LOOP_RESTARTING) {
    restart_after(p, g, err);
    return;
  }
  g.err = err;
  g.accepted = 0;
  g.nsteps += 1;
  g.counts[FODE_C_STEPS] += 1;
  if (isnan(err) || isinf(err)) {
    g.status = FODE_NAN_OR_INF;
    ctrl_finish(g);
    return;
  }
  if (err < 1.0 && g.shortened) {
    g.counts[FODE_C_ACCEPTED] += 1;
    g.t = g.dt;
    g.accepted = 1;
    g.proposed = g.H_ctrl;
    ctrl_outer(p, g);
  } else if (err < 1.0) {
    const double fac = ctrl_factor(p, g, err);
    g.counts[FODE_C_ACCEPTED] += 1;
    g.t = g.last ? g.dt : g.t + g.H;
    g.accepted = 1;
    g.err_prev = err;
    g.h_prev = g.H;
    double Hn = g.H * fac;
    if (g.rej_last) Hn = fmin(Hn, g.H);
    Hn = fmax(g.h_min, fmin(Hn, g.h_max));
    if (p.quantize) Hn = quantize(Hn, g.dt);
    g.rej_last = g.rej_more = 0;
    g.proposed = Hn;
    g.H = Hn;
    ctrl_outer(p, g);
  } else {
    double Hn = g.H * fmax(p.fac_min, p.safety / pow(err, 1.0 / p.k_exp));
    if (g.rej_more) Hn = g.H * 0.1;
    g.rej_more = g.rej_last;
    g.rej_last = 1;
    if (p.quantize) Hn = quantize(Hn, g.dt);
    g.H = Hn;
    const double rem = g.dt - g.t;
    g.last = rem <= g.H * (1.0 + 1.0e-8);
    g.shortened = g.last && rem < g.H * (1.0 - 1.0e-8);
    g.H_ctrl = g.H;
    if (g.last) g.H = rem;
    g.counts[FODE_C_REJECTED] += 1;
    ctrl_inner(p, g, FODE_RETRY);
  }
}

// Workgroup sum of one double per thread (deterministic tree order).
template <int NTH>
__device__ inline double block_sum(double x, double* red) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = x;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < NTH / 64; ++i) s += red[i];
  }
  __syncthreads();
  return s;   // valid in thread 0
}

// One cell's share of error_norm's sum (fastode/solvers.py:37-43):
//   sum_i (E_i / (atol_i + rtol * max(|Y_i|, |Ynew_i|)))^2,
// thread t taking i = t, t + NT, ... and then block_sum's tree.  The plain
// and the batched forms both sum a group cell by cell in this order, so they
// see the same error.
__device__ inline double cell_err_sum(const double* Y, const double* Ynew, const double* E,
                                      const double* atol, double rtol, int n, double* red) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) {
    const double ymax = fmax(fabs(Y[i]), fabs(Ynew[i]));
    const double s = E[i] / (atol[i] + rtol * ymax);
    acc = acc + s * s;
  }
  return block_sum<NT>(acc, red);
}

// The same sum with the algebraic rows (mass 0) of E replaced by the step
// change Ynew - Y (RestartingLoop's step_change_rows, solvers.py:240-241).
__device__ inline double cell_err_sum_dae(const double* Y, const double* Ynew, const double* E,
                                          const double* atol, double rtol, int n, double* red,
                                          const double* mass) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) {
    const double ymax = fmax(fabs(Y[i]), fabs(Ynew[i]));
    const double e = mass[i] == 0.0 ? Ynew[i] - Y[i] : E[i];
    const double s = e / (atol[i] + rtol * ymax);
    acc = acc + s * s;
  }
  return block_sum<NT>(acc, red);
}

// Whether the step-change rows apply to an attempt of this solver.
__host__ __device__ inline bool step_change_rows(const fode_ctrl_params& p) {
  return p.loop == FODE_LOOP_RESTARTING && p.step_change_rows && p.constrained;
}

// error_norm's value from the group's sum over size entries.
__device__ inline double err_from_sum(double sum, double size) {
  const double v = sqrt(sum / size);
  return isnan(v) ? v : fmax(v, 1.0e-10);
}


}  // namespace fode
