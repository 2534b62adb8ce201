// The batched persistent kernel and its launchers, shared by the two
// translation units that instantiate it (tables in LDS: batched.hip; tables
// in memory: batched_unstaged.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "cell.hpp"

// Shared with fastode_hip.hip, which owns the handles and the error text.
extern "C" const void* fode_internal_plan(void* plan);
extern "C" const void* fode_internal_mech(void* mech);
extern "C" void fode_internal_set_error(const char* msg);

namespace fode {


inline int bfail(const std::string& m) {
  fode_internal_set_error(m.c_str());
  return -1;
}

struct BatchArgs {
  PlanDev P;
  MechDev M;
  fode_ctrl_params p;
  fode_batch_desc b;
  Layout lay;
  // Where a group's scratch rows live in b.work: at its cells' rows
  // (work_by_block 0: work is (C, n), its second half from row C), or in the
  // slot of the workgroup that holds the group (1: group_size * n doubles
  // per workgroup, the second halves from work_half on).
  int work_by_block;
  long work_half;
  int max_blocks;   // > 0: launch at most this many workgroups (work slots)
  unsigned long long* timers;   // phase cycle sums (timed builds), else null
};

// A group's scratch rows (half 0: new states / Newton iterates; half 1: the
// projection's saved state).
__device__ __forceinline__ double* group_work(const BatchArgs& A, int c0, int half) {
  const int n = A.M.n;
  if (A.work_by_block)
    return A.b.work + half * A.work_half + (size_t)blockIdx.x * A.b.group_size * n;
  return A.b.work + ((size_t)half * A.b.C + c0) * n;
}

// Tables in LDS when everything fits in one workgroup's LDS (and the caller
// allows it), else in memory.
inline Layout choose_layout(const PlanDev& P, const MechDev& M, int S, int options = 0) {
  int dev = 0, maxlds = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  const Layout staged = make_layout(P, M, S, STAGE_ALL);
  if (!(options & 1) && (size_t)staged.total * sizeof(double) + sizeof(fode_group_ctrl) <= (size_t)maxlds)
    return staged;
  return make_layout(P, M, S, STAGE_NONE);
}

// One attempt on one cell. Returns the cell's error partial sum (thread 0).
template <int TILE>
__device__ __forceinline__ double cell_attempt(const BatchArgs& A, const Tables& T, const EngineLds& L,
                               const CellLds& V, const fode_group_ctrl& g, int cell, double* ynew_out) {
  const MechDev& M = A.M;
  const int n = M.n, tid = threadIdx.x;
  const double* yg = A.b.Y + (size_t)cell * n;
  const double* kg = A.b.k + (size_t)cell * A.b.k_stride;
  for (int i = tid; i < n; i += NT) V.y[i] = yg[i];
  // one shared row of rate constants (k_stride 0): V.kc was formed once for
  // the launch (the cell computations only read it)
  if (A.b.k_stride != 0)
    for (int j = tid; j < M.R; j += NT) V.kc[j] = kg[j] * A.b.kmul[j];
  __syncthreads();
  FODE_PHASE(PH_LOAD);
  cell_stages<TILE>(A.P, T, M, A.p, A.b.mass, L, V, g.H, g.alpha, ynew_out);
  const double e = step_change_rows(A.p) ? cell_err_sum_dae(V.y, V.w, V.fr, A.b.atol, A.b.rtol, n, V.red, A.b.mass)
                                         : cell_err_sum(V.y, V.w, V.fr, A.b.atol, A.b.rtol, n, V.red);
  FODE_PHASE(PH_COMBINE);
  return e;
}

// Maximum over the workgroup of one double per thread, to every thread.
__device__ inline double block_max_all(double x, double* red) {
  for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  double m = red[0];
  for (int i = 1; i < NT / 64; ++i) m = fmax(m, red[i]);
  __syncthreads();
  return m;
}

// Cell c of the group (state from Y, rate constants) into LDS.
__device__ __forceinline__ void load_cell(const BatchArgs& A, const CellLds& V, int cell) {
  const int n = A.M.n;
  const double* kg = A.b.k + (size_t)cell * A.b.k_stride;
  for (int i = threadIdx.x; i < n; i += NT) V.y[i] = A.b.Y[(size_t)cell * n + i];
  for (int j = threadIdx.x; j < A.M.R; j += NT) V.kc[j] = kg[j] * A.b.kmul[j];
  __syncthreads();
}

// The projection's residual at y (LDS): F_i(y) on algebraic rows, 0 on the
// others (fastode/solvers.py:110-114), into g.
template <int TILE>
__device__ __forceinline__ void residual_into(const BatchArgs& A, const CellLds& V, const double* y, double* g) {
  rates_into<slot3_paths(TILE)>(A.M, y, V.kc, V.rd);
  for (int i = threadIdx.x; i < A.M.n; i += NT) g[i] = A.b.mass[i] == 0.0 ? forcing(A.M, y, V.rd, i) : 0.0;
  __syncthreads();
}

// out = (M - J_alg(V.y))^{-1} rhs: the factorization with algebraic rows
// only (fastode/linalg.py:60-67, alpha = 1) and one solve.
template <int TILE>
__device__ __forceinline__ void alg_solve(const BatchArgs& A, const Tables& T, const EngineLds& L, const CellLds& V,
                          const double* rhs, double* out) {
  const PlanDev& P = A.P;
  const int n = A.M.n;
  jacobian_into<slot3_paths(TILE)>(A.M, V);
  double a[TILE][TILE];
  assemble<TILE>(P, JLds{V.jv}, 1.0, A.b.mass, true, L.lds, a);
  factor_cell<TILE>(P, T, L, a);
  double* vec = L.lds + P.vec;
  for (int q = threadIdx.x; q < P.c + P.mpad; q += NT) vec[q] = q < n ? rhs[P.perm[q]] : 0.0;
  solve_cell<TILE>(P, T, L, a);
  for (int q = threadIdx.x; q < n; q += NT) out[P.perm[q]] = vec[q];
  __syncthreads();
}

__device__ inline double weighted_term(double z, double d, double atol, double rtol) {
  const double sc = atol + rtol * fmax(fabs(z), fabs(z + d));
  return sc > 0.0 ? fabs(d) / sc : (d == 0.0 ? 0.0 : INFINITY);
}

// project (fastode/solvers.py:100-178) for one group, on the device: damped
// Newton on the algebraic rows with group-wide maxima, the factorization
// recomputed cell by cell.  The group's Newton corrections (algebraic rows)
// live in work, its saved state in the second half of work.  Returns true
// when converged; a failure restores the state.
template <int TILE>
__device__ __forceinline__ bool project_group(const BatchArgs& A, const Tables& T, const EngineLds& L, const CellLds& V,
                              fode_group_ctrl& gs, int c0, int cells) {
  const fode_batch_desc& B = A.b;
  const int n = A.M.n, tid = threadIdx.x;
  const double* mass = B.mass;
  double* Yg = B.Y + (size_t)c0 * n;
  double* Dg = group_work(A, c0, 0);
  double* Og = group_work(A, c0, 1);
  constexpr int MAX_UPDATES = 10, MAX_BACKTRACKS = 24;
  constexpr double TOL = 0.1, FACTOR = 0.5, SIGMA = 1.0e-4;
  auto count = [&](int which) {
    if (tid == 0) gs.counts[which] += 1;
  };
  bool saved = false;
  for (int update = 0; update <= MAX_UPDATES; ++update) {
    double gmax = 0.0, bad = 0.0;
    for (int c = 0; c < cells; ++c) {
      load_cell(A, V, c0 + c);
      residual_into<TILE>(A, V, V.y, V.f0);
      for (int i = tid; i < n; i += NT)
        if (mass[i] == 0.0) {
          const double gi = V.f0[i];
          if (!isfinite(gi)) bad = 1.0;
          gmax = fmax(gmax, fabs(gi));
        }
    }
    count(FODE_C_RHS);
    bad = block_max_all(bad, V.red);
    gmax = block_max_all(gmax, V.red);
    if (bad != 0.0) break;
    count(FODE_C_PROJ_PASSES);
    if (gmax == 0.0) return true;
    count(FODE_C_JAC);
    count(FODE_C_LU);
    count(FODE_C_SOLVES);
    double q = 0.0;
    bad = 0.0;
    for (int c = 0; c < cells; ++c) {
      load_cell(A, V, c0 + c);
      residual_into<TILE>(A, V, V.y, V.f0);
      alg_solve<TILE>(A, T, L, V, V.f0, V.fr);
      for (int i = tid; i < n; i += NT)
        if (mass[i] == 0.0) {
          const double d = V.fr[i];
          Dg[(size_t)c * n + i] = d;
          if (!isfinite(d)) bad = 1.0;
          q = fmax(q, weighted_term(V.y[i], d, B.atol[i], B.rtol));
        }
    }
    bad = block_max_all(bad, V.red);
    q = block_max_all(q, V.red);
    if (bad != 0.0) break;
    if (q <= TOL) return true;
    if (update == MAX_UPDATES) break;
    if (!saved) {
      for (long e = tid; e < (long)cells * n; e += NT)
        if (mass[e % n] == 0.0) Og[e] = Yg[e];
      saved = true;
    }
    double step = 1.0, taken = -1.0;
    for (int bt = 0; bt <= MAX_BACKTRACKS; ++bt) {
      const double used = step;
      step *= FACTOR;
      double cbad = 0.0, gbad = 0.0, gcmax = 0.0;
      for (int c = 0; c < cells; ++c) {
        load_cell(A, V, c0 + c);
        for (int i = tid; i < n; i += NT) {
          double v = V.y[i];
          if (mass[i] == 0.0) {
            v = v + used * Dg[(size_t)c * n + i];
            if (!isfinite(v)) cbad = 1.0;
          }
          V.w[i] = v;
        }
        __syncthreads();
        residual_into<TILE>(A, V, V.w, V.f0);
        for (int i = tid; i < n; i += NT)
          if (mass[i] == 0.0) {
            if (!isfinite(V.f0[i])) gbad = 1.0;
            gcmax = fmax(gcmax, fabs(V.f0[i]));
          }
      }
      cbad = block_max_all(cbad, V.red);
      gbad = block_max_all(gbad, V.red);
      gcmax = block_max_all(gcmax, V.red);
      if (cbad != 0.0) continue;
      count(FODE_C_RHS);
      count(FODE_C_PROJ_TRIALS);
      if (gbad != 0.0) continue;
      if (gcmax == 0.0) {
        taken = used;
        break;
      }
      count(FODE_C_SOLVES);
      double qt = 0.0, tbad = 0.0;
      for (int c = 0; c < cells; ++c) {
        load_cell(A, V, c0 + c);
        for (int i = tid; i < n; i += NT)
          V.w[i] = mass[i] == 0.0 ? V.y[i] + used * Dg[(size_t)c * n + i] : V.y[i];
        __syncthreads();
        residual_into<TILE>(A, V, V.w, V.f0);
        alg_solve<TILE>(A, T, L, V, V.f0, V.fr);     // the factorization at Y, as LU.solve reuses it
        for (int i = tid; i < n; i += NT)
          if (mass[i] == 0.0) {
            const double d = V.fr[i];
            if (!isfinite(d)) tbad = 1.0;
            qt = fmax(qt, weighted_term(V.w[i], d, B.atol[i], B.rtol));
          }
      }
      tbad = block_max_all(tbad, V.red);
      qt = block_max_all(qt, V.red);
      if (tbad != 0.0) continue;
      if (qt <= TOL || qt < (1.0 - SIGMA * used) * q) {
        taken = used;
        break;
      }
    }
    if (taken < 0.0) break;
    for (long e = tid; e < (long)cells * n; e += NT)
      if (mass[e % n] == 0.0) Yg[e] = Yg[e] + taken * Dg[e];
    __syncthreads();
  }
  if (saved) {
    for (long e = tid; e < (long)cells * n; e += NT)
      if (mass[e % n] == 0.0) Yg[e] = Og[e];
  }
  __syncthreads();
  return false;
}

// DAE: the tableau is constrained (projection at the start of each call);
// a separate instantiation keeps the projection's registers out of the
// unconstrained kernel.
// STAGE: which tables are in LDS (copied at the start): none, all, or the
// solves' only; separate instantiations, so that the compiler addresses
// them as LDS rather than through generic pointers.
// WPE: waves per SIMD the registers are held to (1: as many registers as the
// code wants, one wave per SIMD; 2: at most 256 per lane, so that two
// workgroups share a CU when their LDS fits).
template <int STAGE>
__device__ __forceinline__ Tables kernel_tables(const PlanDev& P, double* smem) {
  if constexpr (STAGE == STAGE_ALL) return stage_tables(P, reinterpret_cast<uint32_t*>(smem));
  if constexpr (STAGE == STAGE_SOLVE) return stage_solve_tables(P, reinterpret_cast<uint32_t*>(smem));
  return P.t;
}

template <int TILE, bool DAE, int STAGE, int WPE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_batched(BatchArgs A) {
  extern __shared__ double smem[];
  const int n = A.M.n, S = A.p.stages;
  const Tables T = kernel_tables<STAGE>(A.P, smem);
  const EngineLds L = carve(smem + A.lay.engine, A.P);
  const CellLds V = carve_cell(smem, A.lay, n, A.M.R, S);
  __shared__ fode_group_ctrl gs;
  const int tid = threadIdx.x;
  const fode_batch_desc& B = A.b;
  phase_start();
  if (B.k_stride == 0)
    for (int j = tid; j < A.M.R; j += NT) V.kc[j] = B.k[j] * B.kmul[j];
  for (;;) {
    if (tid == 0) {
      const int idx = atomicAdd(B.queue, 1);
      int grp = idx;
      if (idx < B.ngroups && B.order) {
        grp = B.order[idx];
        if (grp < 0 || grp >= B.ngroups) grp = idx;   // not a permutation: ignore it
      }
      V.flags[0] = grp;
    }
    __syncthreads();
    const int group = V.flags[0];
    __syncthreads();
    if (group >= B.ngroups) break;
    const int c0 = group * B.group_size;
    const int cells = min(B.group_size, B.C - c0);
    if (tid == 0) {
      gs = B.g[group];
      gs.dt = B.dt;
      ctrl_init(A.p, gs);
    }
    __syncthreads();
    // consistent initialization of constrained systems (solvers.py:322-325):
    // a failure returns the restored state, unclamped, with Memory untouched
    bool consistent = true;
    if constexpr (DAE) consistent = project_group<TILE>(A, T, L, V, gs, c0, cells);
    if (tid == 0) {
      if (consistent) {
        ctrl_outer(A.p, gs);
      } else {
        gs.status = FODE_PROJECTION_FAILED;
        gs.action = FODE_FINISH;
      }
    }
    __syncthreads();
    int nrec = 0;   // accepted steps recorded (thread 0)
    while (gs.action != FODE_FINISH) {
      double sum = 0.0;
      for (int c = 0; c < cells; ++c) {
        double* ynew = group_work(A, c0, 0) + (size_t)c * n;
        const double cs = cell_attempt<TILE>(A, T, L, V, gs, c0 + c, ynew);
        sum = sum + cs;
      }
      if (tid == 0) {
        const double h_used = gs.H;
        ctrl_after(A.p, gs, err_from_sum(sum, (double)cells * n));
        if (gs.accepted && B.hlog && nrec < B.hcap) B.hlog[(size_t)group * B.hcap + nrec++] = h_used;
      }
      __syncthreads();
      if (gs.accepted) {
        const long N = (long)cells * n;
        double* Y = B.Y + (size_t)c0 * n;
        const double* W = group_work(A, c0, 0);
        for (long i = tid; i < N; i += NT) Y[i] = W[i];
      }
      __syncthreads();
      FODE_PHASE(PH_CONTROL);
    }
    // clamp (fastode/solvers.py:89-97) and record the call's status
    if (gs.status != FODE_PROJECTION_FAILED) {
      const long N = (long)cells * n;
      double* Y = B.Y + (size_t)c0 * n;
      for (long i = tid; i < N; i += NT) {
        if (A.p.constrained && B.mass[i % n] != 1.0) continue;
        Y[i] = np_max0(Y[i]);
      }
    }
    if (tid == 0) {
      if (gs.status != FODE_CONVERGED) {
        if (gs.fails == 0) gs.first_fail = (int)gs.counts[FODE_C_CALLS] - 1;
        gs.fails += 1;
      }
      B.g[group] = gs;
      if (B.nacc) B.nacc[group] = nrec;
    }
    __syncthreads();
    FODE_PHASE(PH_CONTROL);
  }
  phase_flush(A.timers);
}

// np.maximum(a, b): NaN from either side propagates.
__device__ inline double np_max2(double a, double b) {
  if (isnan(a)) return a;
  if (isnan(b)) return b;
  return a >= b ? a : b;
}

// One Newton pass of backward Euler on one cell (fastode/solvers.py:450-461):
// F and J at the iterate (from work), the factorization of I/H - J, the
// correction for F - (Yn1 - Yn)/H, and the new iterate max(Yn1 + delta, 0)
// back to work.  Returns (in thread 0's view after the block reductions) a
// miss flag: some correction non-finite or above its bound
// (BackwardEuler.converged, :433-438), and sets *nonfinite for an iterate
// that is not finite.
template <int TILE>
__device__ __forceinline__ void be_cell_pass(const BatchArgs& A, const Tables& T, const EngineLds& L,
                                             const CellLds& V, const fode_group_ctrl& g, int cell,
                                             double& miss, double& nonfinite) {
  const MechDev& M = A.M;
  const PlanDev& P = A.P;
  const fode_batch_desc& B = A.b;
  const int n = M.n, tid = threadIdx.x;
  double* yn1 = group_work(A, cell - cell % B.group_size, 0) + (size_t)(cell % B.group_size) * n;
  const double* yn = B.Y + (size_t)cell * n;
  const double* kg = B.k + (size_t)cell * B.k_stride;
  for (int i = tid; i < n; i += NT) {
    V.y[i] = yn1[i];
    V.w[i] = yn[i];
  }
  if (B.k_stride != 0)
    for (int j = tid; j < M.R; j += NT) V.kc[j] = kg[j] * B.kmul[j];
  __syncthreads();
  rates_into<slot3_paths(TILE)>(M, V.y, V.kc, V.rd);
  forcing_into<slot3_paths(TILE)>(M, V.y, V.rd, V.f0);
  jacobian_into<slot3_paths(TILE)>(M, V);
  double a[TILE][TILE];
  assemble<TILE>(P, JLds{V.jv}, g.alpha, B.mass, false, L.lds, a);
  factor_cell<TILE>(P, T, L, a);
  double* vec = L.lds + P.vec;
  const double H = g.H;
  for (int q = tid; q < P.c + P.mpad; q += NT) {
    double v = 0.0;
    if (q < n) {
      const int i = P.perm[q];
      const double d = V.y[i] - V.w[i];
      v = V.f0[i] - d / H;
    }
    vec[q] = v;
  }
  solve_cell<TILE>(P, T, L, a);
  for (int q = tid; q < n; q += NT) V.fr[P.perm[q]] = vec[q];
  __syncthreads();
  for (int i = tid; i < n; i += NT) {
    const double d = V.fr[i];
    const double v = np_max0(V.y[i] + d);
    yn1[i] = v;
    const double bound = np_max2(np_max2(A.p.small, B.atol[i]), B.rtol * fabs(v));
    if (!isfinite(d) || !(fabs(d) <= bound)) miss = 1.0;
    if (!isfinite(v)) nonfinite = 1.0;
  }
  __syncthreads();
}

// Batched backward Euler (BackwardEuler.solve, fastode/solvers.py:440-491)
// on the persistent workgroups: per group, the step's start Yn in Y and the
// Newton iterate Yn1 in work; each pass runs cell by cell and the group's
// decision (be_after) follows the block reductions of the flags.
template <int TILE, int STAGE, int WPE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_batched_be(BatchArgs A) {
  extern __shared__ double smem[];
  const int n = A.M.n, S = A.p.stages;
  const Tables T = kernel_tables<STAGE>(A.P, smem);
  const EngineLds L = carve(smem + A.lay.engine, A.P);
  const CellLds V = carve_cell(smem, A.lay, n, A.M.R, S);
  __shared__ fode_group_ctrl gs;
  const int tid = threadIdx.x;
  const fode_batch_desc& B = A.b;
  if (B.k_stride == 0)
    for (int j = tid; j < A.M.R; j += NT) V.kc[j] = B.k[j] * B.kmul[j];
  for (;;) {
    if (tid == 0) {
      const int idx = atomicAdd(B.queue, 1);
      int grp = idx;
      if (idx < B.ngroups && B.order) {
        grp = B.order[idx];
        if (grp < 0 || grp >= B.ngroups) grp = idx;
      }
      V.flags[0] = grp;
    }
    __syncthreads();
    const int group = V.flags[0];
    __syncthreads();
    if (group >= B.ngroups) break;
    const int c0 = group * B.group_size;
    const int cells = min(B.group_size, B.C - c0);
    const long N = (long)cells * n;
    double* Y = B.Y + (size_t)c0 * n;
    double* W = group_work(A, c0, 0);
    if (tid == 0) {
      gs = B.g[group];
      gs.dt = B.dt;
      ctrl_init(A.p, gs);
    }
    for (long i = tid; i < N; i += NT) W[i] = Y[i];
    __syncthreads();
    while (gs.action != FODE_FINISH) {
      double miss = 0.0, nonfinite = 0.0;
      for (int c = 0; c < cells; ++c) be_cell_pass<TILE>(A, T, L, V, gs, c0 + c, miss, nonfinite);
      miss = block_max_all(miss, V.red);
      nonfinite = block_max_all(nonfinite, V.red);
      if (tid == 0) be_after(A.p, gs, miss == 0.0, nonfinite == 0.0);
      __syncthreads();
      if (gs.accepted == FODE_BE_COMMIT) {
        for (long i = tid; i < N; i += NT) Y[i] = W[i];
      } else if (gs.accepted == FODE_BE_RESTORE) {
        for (long i = tid; i < N; i += NT) W[i] = Y[i];
      }
      __syncthreads();
    }
    // the call returns the clamped iterate (solvers.py:491)
    for (long i = tid; i < N; i += NT) Y[i] = np_max0(W[i]);
    if (tid == 0) {
      if (gs.status != FODE_CONVERGED) {
        if (gs.fails == 0) gs.first_fail = (int)gs.counts[FODE_C_CALLS] - 1;
        gs.fails += 1;
      }
      B.g[group] = gs;
      if (B.nacc) B.nacc[group] = 0;
    }
    __syncthreads();
  }
}

template <int TILE, bool DAE, int STAGE, int WPE>
int launch_batched_as(const BatchArgs& A, int blocks_per_cu, hipStream_t s) {
  const bool be = A.p.loop == FODE_LOOP_BACKWARD_EULER;
  const void* kern = be ? (const void*)k_batched_be<TILE, STAGE, WPE> : (const void*)k_batched<TILE, DAE, STAGE, WPE>;
  const size_t bytes = (size_t)A.lay.total * sizeof(double);
  int dev = 0, cus = 0, maxlds = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (bytes + sizeof(fode_group_ctrl) > (size_t)maxlds)
    return bfail("fode_batched_interval: the plan needs " + std::to_string(bytes) + " bytes of LDS, above " +
                 std::to_string(maxlds));
  hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return bfail(std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  int occ = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, NT, bytes);
  if (e != hipSuccess) (void)hipGetLastError();   // not the launch's error
  if (e != hipSuccess || occ < 1) occ = 1;
  // never more persistent workgroups than can be resident at once
  const int per_cu = blocks_per_cu > 0 && blocks_per_cu < occ ? blocks_per_cu : occ;
  int blocks = cus * per_cu;
  if (blocks > A.b.ngroups) blocks = A.b.ngroups;
  if (A.max_blocks > 0 && blocks > A.max_blocks) blocks = A.max_blocks;
  if (blocks < 1) blocks = 1;
  if (be)
    hipLaunchKernelGGL((k_batched_be<TILE, STAGE, WPE>), dim3(blocks), dim3(NT), bytes, s, A);
  else
    hipLaunchKernelGGL((k_batched<TILE, DAE, STAGE, WPE>), dim3(blocks), dim3(NT), bytes, s, A);
  e = hipGetLastError();
  if (e != hipSuccess) return bfail(std::string("k_batched: ") + hipGetErrorString(e));
  return 0;
}

// The persistent kernel for the plan's tile (after the queue reset).
template <bool STAGED>
int launch_batched_tiles(const BatchArgs& A, int blocks_per_cu, hipStream_t s) {
  // backward Euler has one kernel for both (it takes no projection)
  const bool dae = A.p.constrained != 0 && A.p.loop != FODE_LOOP_BACKWARD_EULER;
  switch (A.P.tile) {
#define X(T)                                                                  \
  case T:                                                                     \
    return dae ? launch_batched_as<T, true, STAGED ? STAGE_ALL : STAGE_NONE, 1>(A, blocks_per_cu, s)   \
               : launch_batched_as<T, false, STAGED ? STAGE_ALL : STAGE_NONE, 1>(A, blocks_per_cu, s);
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10)
#undef X
  }
  return bfail("fode_batched_interval: unsupported tile");
}

// Register tiles up to this size have a two-waves-per-SIMD instantiation.
constexpr int OCC2_MAX_TILE = 5;

// Tables in memory: instantiated in batched_unstaged.hip (its own
// translation unit, compiled in parallel); and the two-waves-per-SIMD
// instantiations (tables in memory) in batched_occ2.hip.
int launch_batched_unstaged(const BatchArgs& A, int blocks_per_cu, hipStream_t s);
int launch_batched_occ2(const BatchArgs& A, int blocks_per_cu, hipStream_t s);

}  // namespace fode
