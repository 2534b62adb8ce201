// Plain-path kernels and the C ABI of the integrator (include/fastode_hip.h).
//
// The plain path keeps the reference's call structure: one launch per
// protocol operation (rhs, jac, factor, solve, stage arithmetic, norm +
// control) on one group of cells, with the controller state on the device.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cell.hpp"

using namespace fode;

// The host binding (the package's _lib.py) lays these structs out with
// ctypes / NumPy records of exactly these sizes.
static_assert(sizeof(fode_group_ctrl) == 240, "fode_group_ctrl layout");
static_assert(sizeof(fode_ctrl_params) == 504, "fode_ctrl_params layout");

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return -1;
}

#define HIP_OK(expr)                                                             \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) return fail(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int upload(const T* host, size_t count, T** dev, std::vector<void*>& owned) {
  *dev = nullptr;
  if (count == 0) return 0;
  HIP_OK(hipMalloc((void**)dev, count * sizeof(T)));
  owned.push_back(*dev);
  HIP_OK(hipMemcpy(*dev, host, count * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// upload, but never empty: a zero element stands in for an empty array, so
// that kernels may read index 0 of a term table unconditionally.
template <class T>
int upload_nonempty(const T* host, size_t count, T** dev, std::vector<void*>& owned) {
  if (count > 0) return upload(host, count, dev, owned);
  const T zero{};
  return upload(&zero, 1, dev, owned);
}

struct MechHandle {
  MechDev d;
  std::vector<void*> owned;
};

struct PlanHandle {
  PlanDev d;
  std::vector<void*> owned;
};

hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ----------------------------------------------------------------- kernels
__global__ void k_rhs(MechDev M, int C, const double* __restrict__ Y, const double* __restrict__ k,
                      double* __restrict__ F) {
  const long total = (long)C * M.n;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int c = (int)(idx / M.n), i = (int)(idx % M.n);
    const double* y = Y + (size_t)c * M.n;
    const double* kc = k + (size_t)c * M.R;
    double f;
    const int ci = M.con_of_row[i];
    if (ci >= 0) {
      f = 0.0;
      const double* cf = M.con_coef + (size_t)ci * M.n;
      for (int s = 0; s < M.n; ++s) f = f + cf[s] * y[s];
      f = f - M.con_total[ci];
    } else {
      // in term order, TERMS terms' rates in flight together
      f = 0.0;
      int q = M.rhs_ptr[i];
      const int qe = M.rhs_ptr[i + 1];
      for (; q + TERMS <= qe; q += TERMS) {
        double r[TERMS], cf[TERMS];
#pragma unroll
        for (int u = 0; u < TERMS; ++u) {
          r[u] = rate(M, y, kc, M.rhs_rxn[q + u]);
          cf[u] = M.rhs_coef[q + u];
        }
#pragma unroll
        for (int u = 0; u < TERMS; ++u) f = f + r[u] * cf[u];
      }
      for (; q < qe; ++q) f = f + rate(M, y, kc, M.rhs_rxn[q]) * M.rhs_coef[q];
    }
    F[idx] = f;
  }
}

__global__ void k_jac(MechDev M, int C, const double* __restrict__ Y, const double* __restrict__ k,
                      double* __restrict__ J) {
  const long total = (long)C * M.nnz;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int c = (int)(idx / M.nnz), e = (int)(idx % M.nnz);
    const double* y = Y + (size_t)c * M.n;
    const double* kc = k + (size_t)c * M.R;
    // in term order, TERMS terms' partials in flight together
    double v = 0.0;
    int q = M.jac_ptr[e];
    const int qe = M.jac_ptr[e + 1];
    for (; q + TERMS <= qe; q += TERMS) {
      double d[TERMS], cf[TERMS];
#pragma unroll
      for (int u = 0; u < TERMS; ++u) {
        const int jp = M.jac_src[q + u];
        d[u] = drate(M, y, kc, jp / M.P, jp % M.P);
        cf[u] = M.jac_coef[q + u];
      }
#pragma unroll
      for (int u = 0; u < TERMS; ++u) v = v + cf[u] * d[u];
    }
    for (; q < qe; ++q) {
      const int jp = M.jac_src[q];
      v = v + M.jac_coef[q] * drate(M, y, kc, jp / M.P, jp % M.P);
    }
    J[idx] = v + M.jac_const[e];
  }
}

// One workgroup per cell: the cell's state and its R rates (k_rhs_cell) or
// R*P rate partials (k_jac_cell) in LDS, each computed once, then every
// forcing row or Jacobian entry summed from LDS in term order.  These are
// the batched kernel's rates_into / forcing_into / jacobian_into, so the
// bits are those of k_rhs / k_jac (the same products, added in the same
// order), without a long row's terms each recomputing their rate from
// global tables.  The host uses them when the LDS fits (cell_lds_bytes);
// k_rhs / k_jac serve larger mechanisms.
__global__ __launch_bounds__(NT) void k_rhs_cell(MechDev M, const double* __restrict__ Y,
                                                 const double* __restrict__ k, double* __restrict__ F) {
  extern __shared__ double smem[];
  const int c = blockIdx.x;
  double* y = smem;
  double* rates = smem + M.n;
  const double* yc = Y + (size_t)c * M.n;
  for (int i = threadIdx.x; i < M.n; i += NT) y[i] = yc[i];
  __syncthreads();
  rates_into(M, y, k + (size_t)c * M.R, rates);
  forcing_into(M, y, rates, F + (size_t)c * M.n);
}

__global__ __launch_bounds__(NT) void k_jac_cell(MechDev M, const double* __restrict__ Y,
                                                 const double* __restrict__ k, double* __restrict__ J) {
  extern __shared__ double smem[];
  const int c = blockIdx.x;
  CellLds V{};
  V.y = smem;
  V.rd = smem + M.n;
  V.kc = const_cast<double*>(k + (size_t)c * M.R);   // read only
  V.jv = J + (size_t)c * M.nnz;
  const double* yc = Y + (size_t)c * M.n;
  for (int i = threadIdx.x; i < M.n; i += NT) V.y[i] = yc[i];
  __syncthreads();
  jacobian_into(M, V);
}

struct JGlobal {
  const double* J;
  __device__ double operator()(int e) const { return J[e]; }
};

// Doubles of LDS ahead of the engine's region when k_solve copies the
// solves' tables there.
__host__ __device__ inline int solve_table_doubles(const PlanDev& P) {
  return (solve_table_words(P.n_sf_tasks, P.sf_nst, P.n_sb_tasks, P.sb_nst, P.n_s_pairs) + 1) / 2;
}

// The factorization reads its tables from memory: copying them (49 KB for
// TS1's plan) into each cell's LDS cost more than it saved (52.1 -> 53.2 us).
template <int TILE>
__global__ __launch_bounds__(NT) void k_factor(PlanDev P, const double* __restrict__ alpha_p,
                                               const double* __restrict__ mass,
                                               const double* __restrict__ J, int alg_only,
                                               double* __restrict__ LU) {
  extern __shared__ double smem[];
  const EngineLds L = carve(smem, P);
  const int cell = blockIdx.x, tid = threadIdx.x;
  const double alpha = *alpha_p;
  double a[TILE][TILE];
  assemble<TILE>(P, JGlobal{J + (size_t)cell * P.nnz}, alpha, mass, alg_only != 0, L.lds, a);
  factor_cell<TILE>(P, P.t, L, a);
  double* out = LU + (size_t)cell * P.lu_len;
  for (int s = tid; s < P.ns; s += NT) out[s] = L.lds[s];
  double* core = out + P.ns;
#pragma unroll
  for (int ti = 0; ti < TILE; ++ti)
#pragma unroll
    for (int tj = 0; tj < TILE; ++tj) core[(ti * TILE + tj) * NT + tid] = a[ti][tj];
  double* dd = core + TILE * TILE * NT;
  for (int i = tid; i < 2 * TILE * NT; i += NT) dd[i] = L.dL[i];   // dL and dU are adjacent
}

// STAGED: the solves' tables are copied into LDS ahead of the engine's
// region (a separate instantiation, so that they are addressed as LDS); the
// arithmetic, and so the bits, are the same.
template <int TILE, bool STAGED>
__global__ __launch_bounds__(NT) void k_solve(PlanDev P, const double* __restrict__ LU,
                                              const double* B, double* X) {
  extern __shared__ double smem[];
  Tables T = P.t;
  int off = 0;
  if constexpr (STAGED) {
    T = stage_solve_tables(P, reinterpret_cast<uint32_t*>(smem));
    off = solve_table_doubles(P);
  }
  const EngineLds L = carve(smem + off, P);
  const int cell = blockIdx.x, tid = threadIdx.x;
  const double* in = LU + (size_t)cell * P.lu_len;
  double a[TILE][TILE];
  for (int s = tid; s < P.ns; s += NT) L.lds[s] = in[s];
  const double* core = in + P.ns;
#pragma unroll
  for (int ti = 0; ti < TILE; ++ti)
#pragma unroll
    for (int tj = 0; tj < TILE; ++tj) a[ti][tj] = core[(ti * TILE + tj) * NT + tid];
  const double* dd = core + TILE * TILE * NT;
  for (int i = tid; i < 2 * TILE * NT; i += NT) L.dL[i] = dd[i];
  double* vec = L.lds + P.vec;
  const double* b = B + (size_t)cell * P.n;
  for (int i = tid; i < P.c + P.mpad; i += NT) vec[i] = (i < P.n) ? b[P.perm[i]] : 0.0;
  solve_cell<TILE>(P, T, L, a);
  double* x = X + (size_t)cell * P.n;
  for (int i = tid; i < P.n; i += NT) x[P.perm[i]] = vec[i];
}

struct KPtrs {
  const double* k[4];
};

__global__ void k_stage_point(fode_ctrl_params p, int r, long N, const double* __restrict__ Y,
                              KPtrs K, double* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    double v = Y[i];
    for (int s = 0; s < r; ++s) {
      const double a = p.a[r * 4 + s];
      if (a != 0.0) v = v + a * K.k[s][i];
    }
    out[i] = v;
  }
}

__global__ void k_stage_forcing(fode_ctrl_params p, int r, long N, int n,
                                const fode_group_ctrl* __restrict__ g, const double* __restrict__ F,
                                KPtrs K, const double* __restrict__ mass, int dae, double* out) {
  const double H = g->H;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    double v = F[i];
    for (int s = 0; s < r; ++s) {
      const double coef = p.c[r * 4 + s] / H;
      if (coef != 0.0) {
        const double ks = dae ? K.k[s][i] * mass[i % n] : K.k[s][i];
        v = v + coef * ks;
      }
    }
    out[i] = v;
  }
}

__global__ void k_stage_final(fode_ctrl_params p, long N, const double* __restrict__ Y, KPtrs K,
                              double* __restrict__ Ynew, double* __restrict__ E) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    double y = Y[i], e = 0.0;
    for (int r = 0; r < p.stages; ++r) {
      const double kr = K.k[r][i];
      if (p.m[r] != 0.0) y = y + p.m[r] * kr;
      if (p.e[r] != 0.0) e = e + p.e[r] * kr;
    }
    Ynew[i] = y;
    E[i] = e;
  }
}

__device__ inline void clamp_range(double* Y, long N, int n, const double* mass, int constrained,
                                   long start, long stride) {
  for (long i = start; i < N; i += stride) {
    if (constrained && mass[i % n] != 1.0) continue;
    Y[i] = np_max0(Y[i]);
  }
}

__global__ void k_ctrl_init(fode_ctrl_params p, fode_group_ctrl* g, int ngroups, int begin_later) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ngroups) return;
  fode_group_ctrl c = g[i];
  ctrl_init(p, c);
  if (!begin_later) ctrl_outer(p, c);
  g[i] = c;
}

__global__ void k_ctrl_begin(fode_ctrl_params p, fode_group_ctrl* g, int ngroups) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ngroups) return;
  fode_group_ctrl c = g[i];
  ctrl_outer(p, c);
  g[i] = c;
}

__global__ __launch_bounds__(256) void k_norm_control(fode_ctrl_params p, fode_group_ctrl* g, int C,
                                                      int n, double* Y, const double* Ynew,
                                                      const double* E, const double* atol,
                                                      double rtol, const double* mass) {
  // The group's sum, cell by cell, each cell's share formed as cell_err_sum
  // forms it (thread t's entries t, t + NT, ..., its wave's xor-shuffle tree,
  // then the wave sums in order), but without its barriers: every wave keeps
  // its sum of each cell of a chunk in part, and thread 0 adds them up in the
  // same order.  Same bits as the batched kernel's cell_err_sum.
  constexpr int CHUNK = 64, WAVES = NT / 64;
  __shared__ double part[CHUNK * WAVES];
  __shared__ int flags[2];
  const long N = (long)C * n;
  double sum = 0.0;
  const bool step_rows = step_change_rows(p);
  constexpr int BATCH = 8;   // cells whose loads are in flight together
  for (int c0 = 0; c0 < C; c0 += CHUNK) {
    const int ce = min(C, c0 + CHUNK);
    for (int cb = c0; cb < ce; cb += BATCH) {
      double acc[BATCH];
#pragma unroll
      for (int u = 0; u < BATCH; ++u) acc[u] = 0.0;
      for (int i = threadIdx.x; i < n; i += NT) {
        const bool alg = step_rows && mass[i] == 0.0;
        const double sc = atol[i];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
          if (cb + u < ce) {
            const size_t o = (size_t)(cb + u) * n + i;
            const double y = Y[o], yn = Ynew[o];
            const double e = alg ? yn - y : E[o];
            const double s = e / (sc + rtol * fmax(fabs(y), fabs(yn)));
            acc[u] = acc[u] + s * s;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < BATCH; ++u) {
        double x = acc[u];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (cb + u < ce && (threadIdx.x & 63) == 0) part[(cb + u - c0) * WAVES + (threadIdx.x >> 6)] = x;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int c = c0; c < ce; ++c) {
        double s = 0.0;
        for (int w = 0; w < WAVES; ++w) s += part[(c - c0) * WAVES + w];
        sum = sum + s;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double err = err_from_sum(sum, (double)N);
    fode_group_ctrl c = *g;
    ctrl_after(p, c, err);
    *g = c;
    flags[0] = c.accepted;
    flags[1] = c.action;
  }
  __syncthreads();
  if (flags[0]) {
    // eight loads in flight before their stores (Y and Ynew are distinct arrays)
    const long step = blockDim.x;
    long i = threadIdx.x;
    for (; i + 7 * step < N; i += 8 * step) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = Ynew[i + u * step];
#pragma unroll
      for (int u = 0; u < 8; ++u) Y[i + u * step] = v[u];
    }
    for (; i < N; i += step) Y[i] = Ynew[i];
  }
  if (flags[1] == FODE_FINISH) {
    __syncthreads();
    clamp_range(Y, N, n, mass, p.constrained, threadIdx.x, blockDim.x);
  }
}

// The same norm and controller over three launches, for groups whose cells
// are many enough that one workgroup's loads in flight bound the kernel above.
// k_norm_cells: workgroup c forms cell c's share exactly as k_norm_control
// forms it (thread t's entries t, t + NT, ..., the wave's xor-shuffle tree,
// the wave sums in order) into share[c].
__global__ __launch_bounds__(256) void k_norm_cells(fode_ctrl_params p, int n, const double* __restrict__ Y,
                                                    const double* __restrict__ Ynew, const double* __restrict__ E,
                                                    const double* __restrict__ atol, double rtol,
                                                    const double* __restrict__ mass, double* __restrict__ share) {
  constexpr int WAVES = NT / 64;
  __shared__ double part[WAVES];
  const int c = blockIdx.x;
  const bool step_rows = step_change_rows(p);
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) {
    const size_t o = (size_t)c * n + i;
    const bool alg = step_rows && mass[i] == 0.0;
    const double y = Y[o], yn = Ynew[o];
    const double e = alg ? yn - y : E[o];
    const double s = e / (atol[i] + rtol * fmax(fabs(y), fabs(yn)));
    acc = acc + s * s;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < WAVES; ++w) s += part[w];
    share[c] = s;
  }
}

// The cells' shares added in cell order, then the controller (one thread).
__global__ void k_norm_decide(fode_ctrl_params p, fode_group_ctrl* g, int C, int n, const double* __restrict__ share) {
  if (threadIdx.x != 0) return;
  double sum = 0.0;
  for (int c = 0; c < C; ++c) sum = sum + share[c];
  fode_group_ctrl gc = *g;
  ctrl_after(p, gc, err_from_sum(sum, (double)C * n));
  *g = gc;
}

// On acceptance Y <- Ynew; on finishing, the clamp of the result.
__global__ void k_norm_apply(long N, int n, const fode_group_ctrl* __restrict__ g, double* __restrict__ Y,
                             const double* __restrict__ Ynew, const double* __restrict__ mass, int constrained) {
  const bool accepted = g->accepted != 0, finish = g->action == FODE_FINISH;
  if (!accepted && !finish) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    double v = accepted ? Ynew[i] : Y[i];
    if (finish && !(constrained && mass[i % n] != 1.0)) v = np_max0(v);
    Y[i] = v;
  }
}

// Backward Euler's Newton right-hand side F - (Yn1 - Yn) / H
// (fastode/solvers.py:456), H from the group's control block.
__global__ void k_be_residual(long N, const fode_group_ctrl* __restrict__ g, const double* __restrict__ F,
                              const double* __restrict__ Yn1, const double* __restrict__ Yn,
                              double* __restrict__ out) {
  const double H = g->H;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const double d = Yn1[i] - Yn[i];
    out[i] = F[i] - d / H;
  }
}

// np.maximum(a, b) of two doubles: NaN from either side propagates.
__device__ inline double np_max(double a, double b) {
  if (isnan(a)) return a;
  if (isnan(b)) return b;
  return a >= b ? a : b;
}

// The rest of a Newton pass over the group (fastode/solvers.py:457-461):
// Yn1 = max(Yn1 + delta, 0), the convergence test over every entry
// (BackwardEuler.converged, :433-438), then the step decision (be_after);
// the iterate becomes the step's start or is reset to it, and at the end of
// the call it is clamped (:491).  One workgroup.
__global__ __launch_bounds__(256) void k_be_update(fode_ctrl_params p, fode_group_ctrl* g, int C, int n,
                                                   double* Yn1, double* Yn, const double* delta,
                                                   const double* atol, double rtol) {
  __shared__ double red[4];
  __shared__ int flags[2];
  const long N = (long)C * n;
  double miss = 0.0, nonfinite = 0.0;
  for (long i = threadIdx.x; i < N; i += blockDim.x) {
    const double d = delta[i];
    const double v = np_max0(Yn1[i] + d);
    Yn1[i] = v;
    const double bound = np_max(np_max(p.small, atol[i % n]), rtol * fabs(v));
    if (!isfinite(d) || !(fabs(d) <= bound)) miss = 1.0;
    if (!isfinite(v)) nonfinite = 1.0;
  }
  for (int off = 32; off > 0; off >>= 1) {
    miss = fmax(miss, __shfl_xor(miss, off, 64));
    nonfinite = fmax(nonfinite, __shfl_xor(nonfinite, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = miss;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < 4; ++w) m = fmax(m, red[w]);
    miss = m;
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nonfinite;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < 4; ++w) m = fmax(m, red[w]);
    fode_group_ctrl c = *g;
    be_after(p, c, miss == 0.0, m == 0.0);
    *g = c;
    flags[0] = c.accepted;
    flags[1] = c.action;
  }
  __syncthreads();
  if (flags[0] == FODE_BE_COMMIT) {
    for (long i = threadIdx.x; i < N; i += blockDim.x) Yn[i] = Yn1[i];
  } else if (flags[0] == FODE_BE_RESTORE) {
    for (long i = threadIdx.x; i < N; i += blockDim.x) Yn1[i] = Yn[i];
  }
  if (flags[1] == FODE_FINISH) {
    __syncthreads();
    for (long i = threadIdx.x; i < N; i += blockDim.x) Yn1[i] = np_max0(Yn1[i]);
  }
}

// The same update over three launches (work: 2 doubles per cell), for
// groups whose cells are many enough that one workgroup's loads bound the
// kernel above.  k_be_cells: workgroup c updates cell c's iterate and
// writes its two flags (a correction outside its bound, a value not
// finite); the group's flags are their maxima, so their order is free.
__global__ __launch_bounds__(256) void k_be_cells(fode_ctrl_params p, int n, double* __restrict__ Yn1,
                                                  const double* __restrict__ delta, const double* __restrict__ atol,
                                                  double rtol, double* __restrict__ flags) {
  __shared__ double red[2 * (NT / 64)];
  const int c = blockIdx.x;
  double miss = 0.0, nonfinite = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) {
    const size_t o = (size_t)c * n + i;
    const double d = delta[o];
    const double v = np_max0(Yn1[o] + d);
    Yn1[o] = v;
    const double bound = np_max(np_max(p.small, atol[i]), rtol * fabs(v));
    if (!isfinite(d) || !(fabs(d) <= bound)) miss = 1.0;
    if (!isfinite(v)) nonfinite = 1.0;
  }
  for (int off = 32; off > 0; off >>= 1) {
    miss = fmax(miss, __shfl_xor(miss, off, 64));
    nonfinite = fmax(nonfinite, __shfl_xor(nonfinite, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    red[2 * (threadIdx.x >> 6)] = miss;
    red[2 * (threadIdx.x >> 6) + 1] = nonfinite;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NT / 64; ++w) {
      miss = fmax(miss, red[2 * w]);
      nonfinite = fmax(nonfinite, red[2 * w + 1]);
    }
    flags[2 * c] = miss;
    flags[2 * c + 1] = nonfinite;
  }
}

// The group's flags, then be_after (one thread).
__global__ void k_be_decide(fode_ctrl_params p, fode_group_ctrl* g, int C, const double* __restrict__ flags) {
  if (threadIdx.x != 0) return;
  double miss = 0.0, nonfinite = 0.0;
  for (int c = 0; c < C; ++c) {
    miss = fmax(miss, flags[2 * c]);
    nonfinite = fmax(nonfinite, flags[2 * c + 1]);
  }
  fode_group_ctrl gc = *g;
  be_after(p, gc, miss == 0.0, nonfinite == 0.0);
  *g = gc;
}

// The iterate committed to Yn or reset to it, then clamped when the call ends.
__global__ void k_be_apply(long N, const fode_group_ctrl* __restrict__ g, double* __restrict__ Yn1,
                           double* __restrict__ Yn) {
  const int keep = g->accepted;
  const bool finish = g->action == FODE_FINISH;
  if (keep == FODE_BE_KEEP && !finish) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    double v = Yn1[i];
    if (keep == FODE_BE_COMMIT) Yn[i] = v;
    else if (keep == FODE_BE_RESTORE) v = Yn[i];
    if (finish) v = np_max0(v);
    Yn1[i] = v;
  }
}

// k[c, cols[q]] = (base[c, q] * f) * scale[q]: the photolysis rewrite at an
// interval start, in the reference's order of products (fastode/ts1.py:205-213).
__global__ void k_scale_columns(int C, int ldk, double* __restrict__ k, int ncols, const int* __restrict__ cols,
                                const double* __restrict__ base, const double* __restrict__ scale,
                                const double* __restrict__ f) {
  const long total = (long)C * ncols;
  const double fv = *f;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int c = (int)(idx / ncols), q = (int)(idx % ncols);
    const double v = base[idx] * fv;
    k[(size_t)c * ldk + cols[q]] = v * scale[q];
  }
}

__global__ void k_clamp(long N, int n, double* Y, const double* mass, int constrained) {
  clamp_range(Y, N, n, mass, constrained, blockIdx.x * (long)blockDim.x + threadIdx.x,
              (long)gridDim.x * blockDim.x);
}

// ---- project's arrays (fastode/solvers.py:100-178), plain path.  The Newton
// decisions stay on the host, one scalar read per decision, as the reference's
// loop takes them; the arrays they test are formed and reduced here.

// out = where(alg, G, 0): the residual with the differential rows exactly 0
// (fastode/solvers.py:124).
__global__ void k_proj_mask(long N, int n, const double* __restrict__ G, const double* __restrict__ mass,
                            double* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    out[i] = mass[i % n] == 0.0 ? G[i] : 0.0;
}

// out = Y with Y + step * D on the algebraic rows: a backtracking candidate
// (fastode/solvers.py:151-153), product then sum, as NumPy rounds them.
__global__ void k_proj_candidate(long N, int n, const double* __restrict__ Y, const double* __restrict__ D, double step,
                                 const double* __restrict__ mass, double* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const double y = Y[i];
    if (mass[i % n] == 0.0) {
      const double sd = step * D[i];
      out[i] = y + sd;
    } else {
      out[i] = y;
    }
  }
}

// Over the algebraic rows of V: out[0] = 1 when an entry is not finite, else
// 0; out[1] = max |V| (Z null: the residual test, fastode/solvers.py:129-132)
// or the weighted max of project's `weighted(Z, V)` (fastode/solvers.py:112-118):
// sc = atol + rtol * maximum(|z|, |z + d|), q = |d| / sc where sc > 0, else 0
// for d == 0 and inf otherwise.  One workgroup; a max does not depend on order.
constexpr int PROJ_NT = 1024;
__global__ void __launch_bounds__(PROJ_NT) k_proj_measure(long N, int n, const double* __restrict__ V,
                                                          const double* __restrict__ Z,
                                                          const double* __restrict__ mass,
                                                          const double* __restrict__ atol, double rtol,
                                                          double* __restrict__ out) {
  __shared__ double s_bad[PROJ_NT / 64], s_max[PROJ_NT / 64];
  double bad = 0.0, m = 0.0;
  for (long i = threadIdx.x; i < N; i += PROJ_NT) {
    const int col = (int)(i % n);
    if (mass[col] != 0.0) continue;
    const double d = V[i];
    if (!isfinite(d)) {
      bad = 1.0;
      continue;
    }
    double q;
    if (Z) {
      const double z = Z[i], a = fabs(z), b = fabs(z + d);
      const double big = isnan(a) ? a : (isnan(b) ? b : (a >= b ? a : b));   // np.maximum
      const double rb = rtol * big;
      const double sc = atol[col] + rb;
      q = sc > 0.0 ? fabs(d) / sc : (d == 0.0 ? 0.0 : INFINITY);
    } else {
      q = fabs(d);
    }
    m = fmax(m, q);
  }
  for (int off = 32; off > 0; off >>= 1) {
    bad = fmax(bad, __shfl_xor(bad, off));
    m = fmax(m, __shfl_xor(m, off));
  }
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
    s_bad[w] = bad;
    s_max[w] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < PROJ_NT / 64; ++k) {
      bad = fmax(bad, s_bad[k]);
      m = fmax(m, s_max[k]);
    }
    out[0] = bad;
    out[1] = m;
  }
}

int grid_for(long N, int block = 256) {
  long g = (N + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

// LDS of k_rhs_cell / k_jac_cell: the state and one value per rate (or
// rate partial), plus one, which an entry without terms reads (its product
// is discarded) when the mechanism has no reactions; within the 64 KB a
// workgroup gets without opting in.
constexpr size_t CELL_LDS_MAX = 65536;
size_t cell_lds_bytes(int n, int values) { return ((size_t)n + (size_t)values + 1) * sizeof(double); }

size_t engine_bytes(const PlanDev& P) {
  return (size_t)engine_lds_doubles(P) * sizeof(double);
}

// Whether bytes of dynamic LDS fit one workgroup of the device.
bool fits_workgroup(size_t bytes) {
  int dev = 0, maxlds = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  return bytes <= (size_t)maxlds;
}

template <int TILE>
int launch_factor(const PlanDev& P, int C, const double* alpha, const double* mass, const double* J,
                  int alg_only, double* LU, hipStream_t s) {
  const size_t bytes = engine_bytes(P);
  HIP_OK(hipFuncSetAttribute((const void*)k_factor<TILE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes));
  hipLaunchKernelGGL(k_factor<TILE>, dim3(C), dim3(NT), bytes, s, P, alpha, mass, J, alg_only, LU);
  HIP_OK(hipGetLastError());
  return 0;
}

// k_solve (one workgroup per cell) copies the solves' tables into LDS
// whenever they fit beside the engine.
template <int TILE>
int launch_solve(const PlanDev& P, int C, const double* LU, const double* B, double* X, hipStream_t s) {
  const size_t staged = engine_bytes(P) + (size_t)solve_table_doubles(P) * sizeof(double);
  if (fits_workgroup(staged)) {
    HIP_OK(hipFuncSetAttribute((const void*)k_solve<TILE, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)staged));
    hipLaunchKernelGGL((k_solve<TILE, true>), dim3(C), dim3(NT), staged, s, P, LU, B, X);
  } else {
    const size_t bytes = engine_bytes(P);
    HIP_OK(hipFuncSetAttribute((const void*)k_solve<TILE, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
    hipLaunchKernelGGL((k_solve<TILE, false>), dim3(C), dim3(NT), bytes, s, P, LU, B, X);
  }
  HIP_OK(hipGetLastError());
  return 0;
}

#define FODE_TILES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10)

KPtrs kptrs(const double* const* K, int count) {
  KPtrs k{};
  for (int i = 0; i < 4; ++i) k.k[i] = (i < count && K) ? K[i] : nullptr;
  return k;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

const char* fode_last_error(void) { return g_err.c_str(); }

// Internal accessors for the other translation units of the library.
const void* fode_internal_plan(void* plan) { return &static_cast<PlanHandle*>(plan)->d; }
const void* fode_internal_mech(void* mech) { return &static_cast<MechHandle*>(mech)->d; }
void fode_internal_set_error(const char* msg) { g_err = msg; }

int fode_device_info(int device, int* lds_per_block, int* cus, int* warp_size) {
  HIP_OK(hipDeviceGetAttribute(lds_per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  HIP_OK(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, device));
  HIP_OK(hipDeviceGetAttribute(warp_size, hipDeviceAttributeWarpSize, device));
  return 0;
}

int fode_mech_create(const fode_mech_desc* d, void** out) {
  if (!d || !out) return fail("fode_mech_create: null argument");
  if (d->n <= 0 || d->R < 0 || d->P <= 0 || d->nnz < d->n) return fail("fode_mech_create: bad sizes");
  if (!d->react || !d->rhs_ptr || !d->jac_ptr || !d->jac_const || (d->ncon > 0 && (!d->con_rows || !d->con_coef ||
                                                                                  !d->con_total)))
    return fail("fode_mech_create: null table");
  // Every index the kernels follow stays inside its array: reactant slots in
  // [0, n] (n pads), forcing terms' reactions in [0, R), Jacobian terms'
  // rate partials in [0, R*P), monotone row pointers.
  for (long jp = 0; jp < (long)d->R * d->P; ++jp)
    if (d->react[jp] < 0 || d->react[jp] > d->n) return fail("fode_mech_create: bad react entry");
  if (d->rhs_ptr[0] != 0 || d->jac_ptr[0] != 0) return fail("fode_mech_create: row pointers must start at 0");
  for (int i = 0; i < d->n; ++i)
    if (d->rhs_ptr[i + 1] < d->rhs_ptr[i]) return fail("fode_mech_create: rhs_ptr not monotone");
  for (int e = 0; e < d->nnz; ++e)
    if (d->jac_ptr[e + 1] < d->jac_ptr[e]) return fail("fode_mech_create: jac_ptr not monotone");
  if ((d->rhs_ptr[d->n] > 0 && (!d->rhs_rxn || !d->rhs_coef)) || (d->jac_ptr[d->nnz] > 0 && (!d->jac_src ||
                                                                                            !d->jac_coef)))
    return fail("fode_mech_create: null table");
  for (int q = 0; q < d->rhs_ptr[d->n]; ++q)
    if (d->rhs_rxn[q] < 0 || d->rhs_rxn[q] >= d->R) return fail("fode_mech_create: bad rhs_rxn");
  for (int q = 0; q < d->jac_ptr[d->nnz]; ++q)
    if (d->jac_src[q] < 0 || (long)d->jac_src[q] >= (long)d->R * d->P) return fail("fode_mech_create: bad jac_src");
  for (int c = 0; c < d->ncon; ++c)
    if (d->con_rows[c] < 0 || d->con_rows[c] >= d->n) return fail("fode_mech_create: bad constraint row");
  auto* h = new MechHandle();
  MechDev& m = h->d;
  m.n = d->n; m.R = d->R; m.P = d->P; m.nnz = d->nnz; m.ncon = d->ncon;
  int rc = 0;
  rc |= upload(d->react, (size_t)d->R * d->P, (int**)&m.react, h->owned);
  rc |= upload(d->rhs_ptr, (size_t)d->n + 1, (int**)&m.rhs_ptr, h->owned);
  rc |= upload_nonempty(d->rhs_rxn, (size_t)d->rhs_ptr[d->n], (int**)&m.rhs_rxn, h->owned);
  rc |= upload_nonempty(d->rhs_coef, (size_t)d->rhs_ptr[d->n], (double**)&m.rhs_coef, h->owned);
  rc |= upload(d->jac_ptr, (size_t)d->nnz + 1, (int**)&m.jac_ptr, h->owned);
  rc |= upload_nonempty(d->jac_src, (size_t)d->jac_ptr[d->nnz], (int**)&m.jac_src, h->owned);
  rc |= upload_nonempty(d->jac_coef, (size_t)d->jac_ptr[d->nnz], (double**)&m.jac_coef, h->owned);
  rc |= upload(d->jac_const, (size_t)d->nnz, (double**)&m.jac_const, h->owned);
  std::vector<int> con_of(d->n, -1);
  for (int c = 0; c < d->ncon; ++c) con_of[d->con_rows[c]] = c;
  rc |= upload(con_of.data(), con_of.size(), (int**)&m.con_of_row, h->owned);
  {
    // S^T by reaction from the species-major forcing tables, and the reactant
    // slots of each species, both in ascending order.
    const int nz = d->rhs_ptr[d->n];
    std::vector<int> st_ptr(d->R + 1, 0), st_sp(nz);
    std::vector<double> st_coef(nz);
    for (int q = 0; q < nz; ++q) ++st_ptr[d->rhs_rxn[q] + 1];
    for (int j = 0; j < d->R; ++j) st_ptr[j + 1] += st_ptr[j];
    std::vector<int> fill(st_ptr.begin(), st_ptr.end() - 1);
    for (int i = 0; i < d->n; ++i)
      for (int q = d->rhs_ptr[i]; q < d->rhs_ptr[i + 1]; ++q) {
        const int at = fill[d->rhs_rxn[q]]++;
        st_sp[at] = i;
        st_coef[at] = d->rhs_coef[q];
      }
    std::vector<int> sp_ptr(d->n + 1, 0), sp_jp;
    for (int jp = 0; jp < d->R * d->P; ++jp)
      if (d->react[jp] >= 0 && d->react[jp] < d->n) ++sp_ptr[d->react[jp] + 1];
    for (int i = 0; i < d->n; ++i) sp_ptr[i + 1] += sp_ptr[i];
    sp_jp.resize(sp_ptr[d->n]);
    std::vector<int> sfill(sp_ptr.begin(), sp_ptr.end() - 1);
    for (int jp = 0; jp < d->R * d->P; ++jp)
      if (d->react[jp] >= 0 && d->react[jp] < d->n) sp_jp[sfill[d->react[jp]]++] = jp;
    rc |= upload(st_ptr.data(), st_ptr.size(), (int**)&m.st_ptr, h->owned);
    rc |= upload(st_sp.data(), st_sp.size(), (int**)&m.st_sp, h->owned);
    rc |= upload(st_coef.data(), st_coef.size(), (double**)&m.st_coef, h->owned);
    rc |= upload(sp_ptr.data(), sp_ptr.size(), (int**)&m.sp_ptr, h->owned);
    rc |= upload(sp_jp.data(), sp_jp.size(), (int**)&m.sp_jp, h->owned);
  }
  if (d->ncon > 0) {
    rc |= upload(d->con_rows, (size_t)d->ncon, (int**)&m.con_rows, h->owned);
    rc |= upload(d->con_coef, (size_t)d->ncon * d->n, (double**)&m.con_coef, h->owned);
    rc |= upload(d->con_total, (size_t)d->ncon, (double**)&m.con_total, h->owned);
  }
  if (rc) {
    for (void* p : h->owned) (void)hipFree(p);
    delete h;
    return -1;
  }
  *out = h;
  return 0;
}

int fode_mech_destroy(void* mech) {
  auto* h = static_cast<MechHandle*>(mech);
  if (!h) return 0;
  for (void* p : h->owned) (void)hipFree(p);
  delete h;
  return 0;
}

int fode_plan_create(const fode_plan_desc* d, void** out, int64_t* lu_len) {
  if (!d || !out) return fail("fode_plan_create: null argument");
  if (d->tile < 1 || d->tile > 10) return fail("fode_plan_create: tile must be 1..10");
  if (d->m > G * d->tile || d->c + d->m != d->n) return fail("fode_plan_create: bad core split");
  if (d->slot_stride < d->tile * d->tile || d->slot_stride % 8) return fail("fode_plan_create: bad slot stride");
  auto* h = new PlanHandle();
  PlanDev& p = h->d;
  p.n = d->n; p.c = d->c; p.m = d->m; p.tile = d->tile; p.ns = d->ns; p.nsch = d->nsch; p.nnz = d->nnz;
  p.mpad = G * d->tile;
  p.slot_stride = d->slot_stride;
  p.vec = d->ns;
  p.lds_len = p.vec + d->c + p.mpad;
  p.lu_len = d->ns + p.mpad * p.mpad + 2 * d->tile * NT;
  p.f_nst = d->f_nst; p.sf_nst = d->sf_nst; p.sb_nst = d->sb_nst;
  p.n_f_tasks = d->n_f_tasks; p.n_f_pairs = d->n_f_pairs; p.n_sf_tasks = d->n_sf_tasks;
  p.n_sb_tasks = d->n_sb_tasks; p.n_s_pairs = d->n_s_pairs; p.n_sch_pairs = d->n_sch_pairs;
  int rc = 0;
  rc |= upload(d->sp_src, (size_t)d->ns, (int**)&p.sp_src, h->owned);
  rc |= upload(d->sp_diag, (size_t)d->ns, (int**)&p.sp_diag, h->owned);
  rc |= upload(d->sp_row, (size_t)d->ns, (int**)&p.sp_row, h->owned);
  rc |= upload((const short*)d->core_src, (size_t)NT * d->slot_stride, (short**)&p.core_src, h->owned);
  p.sch_depth = 0;
  for (int t = 0; t < NT; ++t) {
    int k = 0;
    for (int w = 0; w < 4; ++w) k += __builtin_popcount(d->sch_mask[t * 4 + w]);
    if (k > p.sch_depth) p.sch_depth = k;
    if (d->sch_base[t] % 4 || d->sch_base[t + 1] < d->sch_base[t]) {
      delete h;
      return fail("fode_plan_create: Schur pair ranges must be ordered groups of 4");
    }
  }
  rc |= upload(d->sch_mask, (size_t)NT * 4, (uint32_t**)&p.sch_mask, h->owned);
  rc |= upload(d->sch_base, (size_t)NT + 1, (int**)&p.sch_base, h->owned);
  rc |= upload(d->sch_pairs, (size_t)d->n_sch_pairs, (uint32_t**)&p.t.sch_pairs, h->owned);
  rc |= upload(d->core_row, (size_t)p.mpad, (int**)&p.core_row, h->owned);
  rc |= upload((const uint2*)d->f_tasks, (size_t)d->n_f_tasks, (uint2**)&p.t.f_tasks, h->owned);
  rc |= upload(d->f_ptr, (size_t)d->f_nst + 1, (int**)&p.t.f_ptr, h->owned);
  rc |= upload(d->f_pairs, (size_t)d->n_f_pairs, (uint32_t**)&p.t.f_pairs, h->owned);
  rc |= upload((const uint2*)d->sf_tasks, (size_t)d->n_sf_tasks, (uint2**)&p.t.sf_tasks, h->owned);
  rc |= upload(d->sf_ptr, (size_t)d->sf_nst + 1, (int**)&p.t.sf_ptr, h->owned);
  rc |= upload((const uint2*)d->sb_tasks, (size_t)d->n_sb_tasks, (uint2**)&p.t.sb_tasks, h->owned);
  rc |= upload(d->sb_ptr, (size_t)d->sb_nst + 1, (int**)&p.t.sb_ptr, h->owned);
  rc |= upload(d->s_pairs, (size_t)d->n_s_pairs, (uint32_t**)&p.t.s_pairs, h->owned);
  rc |= upload(d->perm, (size_t)d->n, (int**)&p.perm, h->owned);
  if (rc) {
    for (void* q : h->owned) (void)hipFree(q);
    delete h;
    return -1;
  }
  if (lu_len) *lu_len = p.lu_len;
  *out = h;
  return 0;
}

int fode_plan_destroy(void* plan) {
  auto* h = static_cast<PlanHandle*>(plan);
  if (!h) return 0;
  for (void* p : h->owned) (void)hipFree(p);
  delete h;
  return 0;
}

int fode_rhs(void* mech, int C, const double* Y, const double* k, double* F, void* stream) {
  if (!mech || C < 0) return fail("fode_rhs: bad argument");
  if (C == 0) return 0;
  const MechDev& M = static_cast<MechHandle*>(mech)->d;
  const size_t bytes = cell_lds_bytes(M.n, M.R);
  if (bytes <= CELL_LDS_MAX)
    hipLaunchKernelGGL(k_rhs_cell, dim3(C), dim3(NT), bytes, S(stream), M, Y, k, F);
  else
    hipLaunchKernelGGL(k_rhs, dim3(grid_for((long)C * M.n)), dim3(256), 0, S(stream), M, C, Y, k, F);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_jac(void* mech, int C, const double* Y, const double* k, double* J, void* stream) {
  if (!mech || C < 0) return fail("fode_jac: bad argument");
  if (C == 0) return 0;
  const MechDev& M = static_cast<MechHandle*>(mech)->d;
  const size_t bytes = cell_lds_bytes(M.n, M.R * M.P);
  if (bytes <= CELL_LDS_MAX)
    hipLaunchKernelGGL(k_jac_cell, dim3(C), dim3(NT), bytes, S(stream), M, Y, k, J);
  else
    hipLaunchKernelGGL(k_jac, dim3(grid_for((long)C * M.nnz)), dim3(256), 0, S(stream), M, C, Y, k, J);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_factor(void* plan, int C, const double* alpha, const double* mass, const double* J,
                int alg_only, double* LU, void* stream) {
  if (!plan || C < 0 || !alpha) return fail("fode_factor: bad argument");
  if (C == 0) return 0;
  const PlanDev& P = static_cast<PlanHandle*>(plan)->d;
  switch (P.tile) {
#define X(T) case T: return launch_factor<T>(P, C, alpha, mass, J, alg_only, LU, S(stream));
    FODE_TILES(X)
#undef X
  }
  return fail("fode_factor: unsupported tile");
}

int fode_solve(void* plan, int C, const double* LU, const double* B, double* X, void* stream) {
  if (!plan || C < 0) return fail("fode_solve: bad argument");
  if (C == 0) return 0;
  const PlanDev& P = static_cast<PlanHandle*>(plan)->d;
  switch (P.tile) {
#define X(T) case T: return launch_solve<T>(P, C, LU, B, X, S(stream));
    FODE_TILES(X)
#undef X
  }
  return fail("fode_solve: unsupported tile");
}

int fode_ctrl_init(const fode_ctrl_params* p, fode_group_ctrl* g, int ngroups, int begin_later,
                   void* stream) {
  if (!p || !g || ngroups < 0) return fail("fode_ctrl_init: bad argument");
  if (ngroups == 0) return 0;
  hipLaunchKernelGGL(k_ctrl_init, dim3((ngroups + 255) / 256), dim3(256), 0, S(stream), *p, g, ngroups,
                     begin_later);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_ctrl_begin(const fode_ctrl_params* p, fode_group_ctrl* g, int ngroups, void* stream) {
  if (!p || !g || ngroups < 0) return fail("fode_ctrl_begin: bad argument");
  if (ngroups == 0) return 0;
  hipLaunchKernelGGL(k_ctrl_begin, dim3((ngroups + 255) / 256), dim3(256), 0, S(stream), *p, g, ngroups);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_stage_point(const fode_ctrl_params* p, int r, int C, int n, const double* Y,
                     const double* const* K, double* out, void* stream) {
  if (!p || r < 0 || r >= 4) return fail("fode_stage_point: bad argument");
  const long N = (long)C * n;
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_stage_point, dim3(grid_for(N)), dim3(256), 0, S(stream), *p, r, N, Y,
                     kptrs(K, r), out);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_stage_forcing(const fode_ctrl_params* p, int r, int C, int n, const fode_group_ctrl* g,
                       const double* F, const double* const* K, const double* mass, int dae,
                       double* out, void* stream) {
  if (!p || !g || r < 0 || r >= 4) return fail("fode_stage_forcing: bad argument");
  const long N = (long)C * n;
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_stage_forcing, dim3(grid_for(N)), dim3(256), 0, S(stream), *p, r, N, n, g, F,
                     kptrs(K, r), mass, dae, out);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_stage_final(const fode_ctrl_params* p, int C, int n, const double* Y, const double* const* K,
                     double* Ynew, double* E, void* stream) {
  if (!p || p->stages < 1 || p->stages > 4) return fail("fode_stage_final: bad argument");
  const long N = (long)C * n;
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_stage_final, dim3(grid_for(N)), dim3(256), 0, S(stream), *p, N, Y,
                     kptrs(K, p->stages), Ynew, E);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_norm_control(const fode_ctrl_params* p, fode_group_ctrl* g, int C, int n, double* Y,
                      const double* Ynew, const double* E, const double* atol, double rtol,
                      const double* mass, double* work, void* stream) {
  if (!p || !g || C <= 0 || n <= 0) return fail("fode_norm_control: bad argument");
  if (!work) {
    hipLaunchKernelGGL(k_norm_control, dim3(1), dim3(256), 0, S(stream), *p, g, C, n, Y, Ynew, E, atol,
                       rtol, mass);
    HIP_OK(hipGetLastError());
    return 0;
  }
  const long N = (long)C * n;
  hipLaunchKernelGGL(k_norm_cells, dim3(C), dim3(256), 0, S(stream), *p, n, Y, Ynew, E, atol, rtol, mass, work);
  HIP_OK(hipGetLastError());
  hipLaunchKernelGGL(k_norm_decide, dim3(1), dim3(64), 0, S(stream), *p, g, C, n, work);
  HIP_OK(hipGetLastError());
  hipLaunchKernelGGL(k_norm_apply, dim3(grid_for(N)), dim3(256), 0, S(stream), N, n, g, Y, Ynew, mass, p->constrained);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_be_residual(int C, int n, const fode_group_ctrl* g, const double* F, const double* Yn1,
                     const double* Yn, double* out, void* stream) {
  if (!g || C < 0 || n <= 0) return fail("fode_be_residual: bad argument");
  const long N = (long)C * n;
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_be_residual, dim3(grid_for(N)), dim3(256), 0, S(stream), N, g, F, Yn1, Yn, out);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_be_update(const fode_ctrl_params* p, fode_group_ctrl* g, int C, int n, double* Yn1, double* Yn,
                   const double* delta, const double* atol, double rtol, double* work, void* stream) {
  if (!p || !g || C <= 0 || n <= 0) return fail("fode_be_update: bad argument");
  if (p->loop != FODE_LOOP_BACKWARD_EULER || p->n_reductions < 0 || p->n_reductions > 8)
    return fail("fode_be_update: parameters are not backward Euler's");
  if (!work) {
    hipLaunchKernelGGL(k_be_update, dim3(1), dim3(256), 0, S(stream), *p, g, C, n, Yn1, Yn, delta, atol, rtol);
    HIP_OK(hipGetLastError());
    return 0;
  }
  const long N = (long)C * n;
  hipLaunchKernelGGL(k_be_cells, dim3(C), dim3(256), 0, S(stream), *p, n, Yn1, delta, atol, rtol, work);
  HIP_OK(hipGetLastError());
  hipLaunchKernelGGL(k_be_decide, dim3(1), dim3(64), 0, S(stream), *p, g, C, work);
  HIP_OK(hipGetLastError());
  hipLaunchKernelGGL(k_be_apply, dim3(grid_for(N)), dim3(256), 0, S(stream), N, g, Yn1, Yn);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_scale_columns(int C, int ldk, double* k, int ncols, const int32_t* cols, const double* base,
                       const double* scale, const double* f, void* stream) {
  if (C < 0 || ncols < 0 || ldk <= 0 || !f) return fail("fode_scale_columns: bad argument");
  const long total = (long)C * ncols;
  if (total == 0) return 0;
  if (!k || !cols || !base || !scale) return fail("fode_scale_columns: null array");
  hipLaunchKernelGGL(k_scale_columns, dim3(grid_for(total)), dim3(256), 0, S(stream), C, ldk, k, ncols, cols, base,
                     scale, f);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_clamp(int C, int n, double* Y, const double* mass, int constrained, void* stream) {
  const long N = (long)C * n;
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_clamp, dim3(grid_for(N)), dim3(256), 0, S(stream), N, n, Y, mass, constrained);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_proj_mask(int C, int n, const double* G, const double* mass, double* out, void* stream) {
  if (C < 0 || n <= 0) return fail("fode_proj_mask: bad shape");
  const long N = (long)C * n;
  if (N == 0) return 0;
  if (!G || !mass || !out) return fail("fode_proj_mask: null array");
  hipLaunchKernelGGL(k_proj_mask, dim3(grid_for(N)), dim3(256), 0, S(stream), N, n, G, mass, out);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_proj_candidate(int C, int n, const double* Y, const double* D, double step, const double* mass, double* out,
                        void* stream) {
  if (C < 0 || n <= 0) return fail("fode_proj_candidate: bad shape");
  const long N = (long)C * n;
  if (N == 0) return 0;
  if (!Y || !D || !mass || !out) return fail("fode_proj_candidate: null array");
  hipLaunchKernelGGL(k_proj_candidate, dim3(grid_for(N)), dim3(256), 0, S(stream), N, n, Y, D, step, mass, out);
  HIP_OK(hipGetLastError());
  return 0;
}

int fode_proj_measure(int C, int n, const double* V, const double* Z, const double* mass, const double* atol,
                      double rtol, double* out, void* stream) {
  if (C < 0 || n <= 0) return fail("fode_proj_measure: bad shape");
  if (!V || !mass || !out || (Z && !atol)) return fail("fode_proj_measure: null array");
  hipLaunchKernelGGL(k_proj_measure, dim3(1), dim3(PROJ_NT), 0, S(stream), (long)C * n, n, V, Z, mass, atol, rtol,
                     out);
  HIP_OK(hipGetLastError());
  return 0;
}

}  // extern "C"
