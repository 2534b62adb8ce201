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

// Top of the step loop (solvers.py:330-348).
__device__ inline void ctrl_outer(const fode_ctrl_params& p, fode_group_ctrl& g) {
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

// Set-up of a call (solvers.py:306-329).
__device__ inline void ctrl_init(const fode_ctrl_params& p, fode_group_ctrl& g) {
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

// error_norm's value from the group's sum over size entries.
__device__ inline double err_from_sum(double sum, double size) {
  const double v = sqrt(sum / size);
  return isnan(v) ? v : fmax(v, 1.0e-10);
}


}  // namespace fode
