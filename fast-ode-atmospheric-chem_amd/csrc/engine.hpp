// Device building blocks of the integrator for gfx950 (wave64, FP64).
//
// The per-cell LU engine: one 256-thread workgroup owns one cell's stage
// matrix alpha*M - J.  The host symbolic plan splits the minimum-degree
// ordered matrix into
//   * sparse entries (min(i,j) < c), held in LDS and computed by
//     level-scheduled dot-product tasks, and
//   * the dense trailing core (m x m, padded to G*TILE with identity), held in
//     registers, 2-D cyclic over a 16 x 16 thread grid: thread (tr, tc) owns
//     core entries (ti*16 + tr, tj*16 + tc), ti, tj < TILE.
// The core is factored right-looking without pivoting with one barrier per
// pivot (double-buffered pivot row/column in LDS); its 16 x 16 diagonal
// blocks are then inverted so that each triangular solve is TILE block steps
// of one barrier each.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fastode_hip.h"

namespace fode {

constexpr int G = 16;
constexpr int NT = G * G;   // threads per cell engine
constexpr double EPS = 2.220446049250313e-16;   // float64 machine epsilon

struct PlanDev {
  int n, c, m, tile, ns, nnz;
  int lds_len, lu_len, mpad;
  const int* sp_src; const int* sp_diag; const int* sp_row;
  const int* core_src; const int* core_diag; const int* core_row;
  const int4* f_tasks; const int* f_ptr; int f_nst; const uint32_t* f_pairs;
  const int* schur_ptr; const uint32_t* schur;
  const int4* sf_tasks; const int* sf_ptr; int sf_nst;
  const int4* sb_tasks; const int* sb_ptr; int sb_nst;
  const uint32_t* s_pairs;
  const int* perm;
};

struct MechDev {
  int n, R, P, nnz, ncon;
  const int* react; const int* rhs_ptr; const int* rhs_rxn; const double* rhs_coef;
  const int* jac_ptr; const int* jac_src; const double* jac_coef; const double* jac_const;
  const int* con_rows; const double* con_coef; const double* con_total;
  const int* con_of_row;   // n: constraint index of a row or -1
};

// LDS map of the cell engine (doubles), sized by the host from the plan.
struct EngineLds {
  double* lds;    // [lds_len]: sparse entries, zero, solve vector (c + mpad)
  double* lbuf;   // [2*mpad]
  double* ubuf;   // [2*mpad]
  double* rbuf;   // [2]
  double* pb;     // [2*NT]
  double* scr;    // [tile*NT]
};

__device__ inline EngineLds carve(double* base, const PlanDev& P) {
  EngineLds e;
  e.lds = base;
  e.lbuf = e.lds + P.lds_len;
  e.ubuf = e.lbuf + 2 * P.mpad;
  e.rbuf = e.ubuf + 2 * P.mpad;
  e.pb = e.rbuf + 2;
  e.scr = e.pb + 2 * NT;
  return e;
}

__host__ __device__ inline int engine_lds_doubles(int lds_len, int mpad, int tile) {
  return lds_len + 4 * mpad + 2 + 2 * NT + tile * NT;
}

// ---------------------------------------------------------------- helpers
__device__ inline double rowsum16(double x) {
  x += __shfl_xor(x, 8, 16);
  x += __shfl_xor(x, 4, 16);
  x += __shfl_xor(x, 2, 16);
  x += __shfl_xor(x, 1, 16);
  return x;
}

// Level-scheduled dot-product tasks over LDS:
//   lds[t] = (lds[t] - sum_q lds[a_q] * lds[b_q]) [/ lds[d]]
// Tasks of one stage are independent; a barrier follows every stage.
__device__ inline void run_tasks(const int4* tasks, const int* ptr, int nst,
                                 const uint32_t* pairs, double* lds) {
  for (int s = 0; s < nst; ++s) {
    const int b = ptr[s], e = ptr[s + 1];
    for (int t = b + (int)threadIdx.x; t < e; t += NT) {
      const int4 tk = tasks[t];
      double acc = lds[tk.x];
      for (int q = tk.y; q < tk.y + tk.z; ++q) {
        const uint32_t w = pairs[q];
        acc = fma(-lds[w >> 16], lds[w & 0xFFFFu], acc);
      }
      if (tk.w >= 0) acc = acc / lds[tk.w];
      lds[tk.x] = acc;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------- assembly
// Value of entry (J index e, old row) of alpha*diag(mass) - J, as
// fastode/linalg.py:24-32 and :60-67 form it: -J, then + alpha*mass on the
// diagonal; alg_only keeps J only on algebraic rows (mass == 0).
template <class JS>
__device__ inline double stage_entry(const JS& jv, int e, int row, int diag, double alpha,
                                     const double* mass, bool alg_only) {
  double v = 0.0;
  if (e >= 0) {
    double j = jv(e);
    if (alg_only && mass[row] != 0.0) j = 0.0;
    v = -j;
  }
  if (diag >= 0) v = v + alpha * mass[diag];
  return v;
}

template <int TILE, class JS>
__device__ inline void assemble(const PlanDev& P, const JS& jv, double alpha, const double* mass,
                                bool alg_only, double* lds, double (&a)[TILE][TILE]) {
  const int tid = threadIdx.x;
  for (int s = tid; s < P.ns; s += NT)
    lds[s] = stage_entry(jv, P.sp_src[s], P.sp_row[s], P.sp_diag[s], alpha, mass, alg_only);
  if (tid == 0) lds[P.ns] = 0.0;
#pragma unroll
  for (int ti = 0; ti < TILE; ++ti)
#pragma unroll
    for (int tj = 0; tj < TILE; ++tj) {
      const int idx = (ti * TILE + tj) * NT + tid;
      const int e = P.core_src[idx];
      a[ti][tj] = (e == -2) ? 1.0
                            : stage_entry(jv, e, P.core_row[idx], P.core_diag[idx], alpha, mass, alg_only);
    }
}

// Schur update of the core from the sparse pivots: per-slot pair lists.
template <int TILE>
__device__ inline void schur(const PlanDev& P, const double* lds, double (&a)[TILE][TILE]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int ti = 0; ti < TILE; ++ti)
#pragma unroll
    for (int tj = 0; tj < TILE; ++tj) {
      const int s = ti * TILE + tj;
      const int b = P.schur_ptr[s], e = P.schur_ptr[s + 1];
      double acc = a[ti][tj];
#pragma nounroll
      for (int q = b; q < e; ++q) {
        const uint32_t w = P.schur[(size_t)q * NT + tid];
        acc = fma(-lds[w >> 16], lds[w & 0xFFFFu], acc);
      }
      a[ti][tj] = acc;
    }
}

// Publish pivot k1 = NB*G + kk1 (column, row, reciprocal pivot) into buffer nb.
template <int TILE, int NB>
__device__ inline void publish(const double (&a)[TILE][TILE], int kk1, int nb, double* lbuf,
                               double* ubuf, double* rbuf, int tr, int tc) {
  constexpr int MP = G * TILE;
  if (tc == kk1) {
#pragma unroll
    for (int ti = NB; ti < TILE; ++ti) lbuf[nb * MP + ti * G + tr] = a[ti][NB];
  }
  if (tr == kk1) {
#pragma unroll
    for (int tj = NB; tj < TILE; ++tj) ubuf[nb * MP + tj * G + tc] = a[NB][tj];
    if (tc == kk1) rbuf[nb] = 1.0 / a[NB][NB];
  }
}

template <int TILE, int B>
__device__ inline void factor_block(double (&a)[TILE][TILE], double* lbuf, double* ubuf,
                                    double* rbuf, int tr, int tc) {
  constexpr int MP = G * TILE;
#pragma nounroll
  for (int kk = 0; kk < G; ++kk) {
    const int k = B * G + kk;
    const int buf = k & 1;
    const double rc = rbuf[buf];
    const double* lb = lbuf + buf * MP + tr;
    const double* ub = ubuf + buf * MP + tc;
    double u[TILE];
#pragma unroll
    for (int tj = B; tj < TILE; ++tj) u[tj] = (tj * G + tc > k) ? ub[tj * G] : 0.0;
    const bool owner = (tc == kk);
#pragma unroll
    for (int ti = B; ti < TILE; ++ti) {
      const bool below = ti * G + tr > k;
      const double l = below ? lb[ti * G] * rc : 0.0;
      if (owner && below) a[ti][B] = l;
#pragma unroll
      for (int tj = B; tj < TILE; ++tj) a[ti][tj] = fma(-l, u[tj], a[ti][tj]);
    }
    if (kk + 1 < G) {
      publish<TILE, B>(a, kk + 1, buf ^ 1, lbuf, ubuf, rbuf, tr, tc);
    } else if constexpr (B + 1 < TILE) {
      publish<TILE, B + 1>(a, 0, buf ^ 1, lbuf, ubuf, rbuf, tr, tc);
    }
    __syncthreads();
  }
}

template <int TILE, int B>
__device__ inline void factor_blocks(double (&a)[TILE][TILE], double* lbuf, double* ubuf,
                                     double* rbuf, int tr, int tc) {
  if constexpr (B < TILE) {
    factor_block<TILE, B>(a, lbuf, ubuf, rbuf, tr, tc);
    factor_blocks<TILE, B + 1>(a, lbuf, ubuf, rbuf, tr, tc);
  }
}

// No-pivot LU of the register-resident core: strictly lower part becomes L
// (unit diagonal implied), the rest U.
template <int TILE>
__device__ inline void core_factor(double (&a)[TILE][TILE], const EngineLds& L) {
  const int tid = threadIdx.x, tr = tid / G, tc = tid % G;
  publish<TILE, 0>(a, 0, 0, L.lbuf, L.ubuf, L.rbuf, tr, tc);
  __syncthreads();
  factor_blocks<TILE, 0>(a, L.lbuf, L.ubuf, L.rbuf, tr, tc);
}

// Inverses of the diagonal 16x16 blocks of L (unit lower) and U; thread
// (tr, tc) keeps entry (tc, tr) of each (the transpose, as the solve uses it).
template <int TILE>
__device__ inline void diag_inverses(const double (&a)[TILE][TILE], double (&dLt)[TILE],
                                     double (&dUt)[TILE], double* scr) {
  const int tid = threadIdx.x, tr = tid / G, tc = tid % G;
  const int task = tid >> 4, col = tid & 15;
  for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
    for (int J = 0; J < TILE; ++J) scr[J * NT + tid] = a[J][J];
    __syncthreads();
    double x[G];
    if (task < TILE) {
      const double* M = scr + task * NT;
      if (pass == 0) {
#pragma unroll
        for (int r = 0; r < G; ++r) {
          double v = (r == col) ? 1.0 : 0.0;
#pragma unroll
          for (int s = 0; s < r; ++s) v = fma(-M[r * G + s], x[s], v);
          x[r] = v;
        }
      } else {
#pragma unroll
        for (int r = G - 1; r >= 0; --r) {
          double v = (r == col) ? 1.0 : 0.0;
#pragma unroll
          for (int s = r + 1; s < G; ++s) v = fma(-M[r * G + s], x[s], v);
          x[r] = v / M[r * G + r];
        }
      }
    }
    __syncthreads();
    if (task < TILE) {
#pragma unroll
      for (int r = 0; r < G; ++r) scr[task * NT + r * G + col] = x[r];
    }
    __syncthreads();
#pragma unroll
    for (int J = 0; J < TILE; ++J) {
      const double d = scr[J * NT + tc * G + tr];
      if (pass == 0) dLt[J] = d; else dUt[J] = d;
    }
    __syncthreads();
  }
}

// Block forward/back substitution of the core: v (length mpad) in LDS
// holds the right-hand side on entry and the solution on exit.
template <int TILE>
__device__ inline void core_solve(const double (&a)[TILE][TILE], const double (&dLt)[TILE],
                                  const double (&dUt)[TILE], double* v, double* pb) {
  const int tid = threadIdx.x, tr = tid / G, tc = tid % G;
  double acc[TILE];
#pragma unroll
  for (int I = 0; I < TILE; ++I) acc[I] = 0.0;
#pragma unroll
  for (int J = 0; J < TILE; ++J) {
    const double s = rowsum16(acc[J]);
    const double r = v[J * G + tr] - s;
    double* p = pb + (J & 1) * NT;
    p[tid] = dLt[J] * r;
    __syncthreads();
    double y = 0.0;
#pragma unroll
    for (int q = 0; q < G; ++q) y += p[q * G + tc];
    if (tr == 0) v[J * G + tc] = y;
#pragma unroll
    for (int I = J + 1; I < TILE; ++I) acc[I] = fma(a[I][J], y, acc[I]);
  }
  __syncthreads();
#pragma unroll
  for (int I = 0; I < TILE; ++I) acc[I] = 0.0;
#pragma unroll
  for (int J = TILE - 1; J >= 0; --J) {
    const double s = rowsum16(acc[J]);
    const double r = v[J * G + tr] - s;
    double* p = pb + (J & 1) * NT;
    p[tid] = dUt[J] * r;
    __syncthreads();
    double x = 0.0;
#pragma unroll
    for (int q = 0; q < G; ++q) x += p[q * G + tc];
    if (tr == 0) v[J * G + tc] = x;
#pragma unroll
    for (int I = 0; I < J; ++I) acc[I] = fma(a[I][J], x, acc[I]);
  }
  __syncthreads();
}

// Full factorization of one cell from assembled values.
template <int TILE>
__device__ inline void factor_cell(const PlanDev& P, const EngineLds& L, double (&a)[TILE][TILE],
                                   double (&dLt)[TILE], double (&dUt)[TILE]) {
  __syncthreads();
  run_tasks(P.f_tasks, P.f_ptr, P.f_nst, P.f_pairs, L.lds);
  schur<TILE>(P, L.lds, a);
  core_factor<TILE>(a, L);
  diag_inverses<TILE>(a, dLt, dUt, L.scr);
}

// Solve with the cell's factor; the right-hand side is already in the LDS
// vector (new ordering, padding zero) and the solution is left there.
template <int TILE>
__device__ inline void solve_cell(const PlanDev& P, const EngineLds& L, const double (&a)[TILE][TILE],
                                  const double (&dLt)[TILE], const double (&dUt)[TILE]) {
  double* vec = L.lds + P.ns + 1;
  __syncthreads();
  run_tasks(P.sf_tasks, P.sf_ptr, P.sf_nst, P.s_pairs, L.lds);
  core_solve<TILE>(a, dLt, dUt, vec + P.c, L.pb);
  run_tasks(P.sb_tasks, P.sb_ptr, P.sb_nst, P.s_pairs, L.lds);
}

// ------------------------------------------------------------ mechanism
// Rate of reaction j: k_j * prod of reactant concentrations (chem.py:86-87,
// the product over the reactant slots first, then times k).
__device__ inline double rate(const MechDev& M, const double* y, const double* k, int j) {
  double x = 1.0;
  bool first = true;
  for (int p = 0; p < M.P; ++p) {
    const int s = M.react[j * M.P + p];
    const double v = (s < M.n) ? y[s] : 1.0;
    x = first ? v : x * v;
    first = false;
  }
  return k[j] * x;
}

// d rate_j / d y_{react[j,p]} = k_j * prod_{q != p} y_{react[j,q]} (chem.py:92-101).
__device__ inline double drate(const MechDev& M, const double* y, const double* k, int j, int p) {
  double prod = k[j];
  for (int q = 0; q < M.P; ++q) {
    if (q == p) continue;
    const int s = M.react[j * M.P + q];
    prod = prod * ((s < M.n) ? y[s] : 1.0);
  }
  return prod;
}

// Forcing of species i given the rates (or, for a constraint row,
// sum_s c_s y_s - total).
__device__ inline double forcing(const MechDev& M, const double* y, const double* rates, int i) {
  const int ci = M.con_of_row[i];
  if (ci >= 0) {
    double g = 0.0;
    const double* cf = M.con_coef + (size_t)ci * M.n;
    for (int s = 0; s < M.n; ++s) g = g + cf[s] * y[s];
    return g - M.con_total[ci];
  }
  double f = 0.0;
  for (int q = M.rhs_ptr[i]; q < M.rhs_ptr[i + 1]; ++q) f = f + rates[M.rhs_rxn[q]] * M.rhs_coef[q];
  return f;
}

// ------------------------------------------------------------ controller
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

// Jacobian value of entry e from the rate partials d (R*P) in LDS, the same
// arithmetic as the standalone Jacobian kernel.
struct JFromD {
  MechDev M;
  const double* d;
  __device__ double operator()(int e) const {
    double v = 0.0;
    for (int q = M.jac_ptr[e]; q < M.jac_ptr[e + 1]; ++q) v = v + M.jac_coef[q] * d[M.jac_src[q]];
    return v + M.jac_const[e];
  }
};

}  // namespace fode
