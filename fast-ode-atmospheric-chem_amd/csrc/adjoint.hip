// Discrete adjoint of the batched form over one host interval.
//
// With the group's accepted steps held fixed (recorded by the forward
// launch), each cell's interval map is a composition of Rosenbrock steps
// (fastode/solvers.py:54-86) and the final clamp (:89-97).  One step from y
// with step H, A = alpha*I - J(y), alpha = 1/(H gamma):
//   w_r = y + sum_{s<r} a_rs K_s,   F_r = F(w_r) (or the last new forcing),
//   A K_r = F_r + (1/H) sum_{s<r} c_rs K_s,   ynew = y + sum_r m_r K_r.
// Its adjoint, stages last first, with Kbar_r = m_r ybarnew and
// ybar = ybarnew to start:
//   mu_r = A^{-T} Kbar_r;  Fbar_src(r) += mu_r;  Kbar_s += (c_rs/H) mu_r;
//   ybar += d/dy [mu_r^T J(y) K_r]  and  kbar_j += (S^T mu_r)_j dr_j/dy . K_r
//   (J is linear in k_j and quadratic-free in y for mass action up to
//   bimolecular products, so the Hessian term is a sum over reactant pairs);
//   for a new forcing at w_r: g = S^T Fbar_r, wbar = sum_j g_j dr_j/dw,
//   kbar_j += g_j r_j(w_r), ybar += wbar, Kbar_s += a_rs wbar.
// A^T is factored by the same LU engine with the plan of the transposed
// pattern under the same ordering; its entries are the same Jacobian values.
#include <hip/hip_runtime.h>

#include <string>

#include "cell.hpp"

using namespace fode;

extern "C" const void* fode_internal_plan(void* plan);
extern "C" const void* fode_internal_mech(void* mech);
extern "C" void fode_internal_set_error(const char* msg);

namespace {

int afail(const std::string& m) {
  fode_internal_set_error(m.c_str());
  return -1;
}

struct AdjArgs {
  PlanDev P, PT;
  MechDev M;
  fode_ctrl_params p;
  fode_adjoint_desc a;
  Layout lay;
  int adj;        // lam, ybar, Kbar, Fbar, kbar
  int phi, gr;    // per-slot scatter values (R*P) and reaction adjoints (R)
  int total;
};

AdjArgs make_args(const PlanDev& P, const PlanDev& PT, const MechDev& M, int S) {
  AdjArgs A;
  A.P = P;
  A.PT = PT;
  A.M = M;
  A.lay = make_layout(P, M, S, STAGE_NONE);
  A.adj = A.lay.total;
  int end = A.adj + 2 * M.n + 2 * S * M.n + M.R;
  const int scr_at = A.lay.engine + engine_scr_offset(P);
  if (M.R * M.P + M.R <= P.tile * NT) {   // the inverse scratch is dead after the factorization
    A.phi = scr_at;
    A.gr = scr_at + M.R * M.P;
  } else {
    A.phi = end;
    A.gr = end + M.R * M.P;
    end += M.R * M.P + M.R;
  }
  A.total = end;
  return A;
}

// (S^T v)_j; four terms' loads in flight per round, the sum in term order
__device__ inline double st_dot(const MechDev& M, const double* v, int j) {
  double g = 0.0;
  int q = M.st_ptr[j];
  const int qe = M.st_ptr[j + 1];
  for (; q + 4 <= qe; q += 4) {
    const int s0 = M.st_sp[q], s1 = M.st_sp[q + 1], s2 = M.st_sp[q + 2], s3 = M.st_sp[q + 3];
    const double c0 = M.st_coef[q], c1 = M.st_coef[q + 1], c2 = M.st_coef[q + 2], c3 = M.st_coef[q + 3];
    const double v0 = v[s0], v1 = v[s1], v2 = v[s2], v3 = v[s3];
    g = g + c0 * v0;
    g = g + c1 * v1;
    g = g + c2 * v2;
    g = g + c3 * v3;
  }
  for (; q < qe; ++q) g = g + M.st_coef[q] * v[M.st_sp[q]];
  return g;
}

// sum of the slot values of species i (same rounds)
__device__ inline double gather_slots(const MechDev& M, const double* phi, int i) {
  double s = 0.0;
  int q = M.sp_ptr[i];
  const int qe = M.sp_ptr[i + 1];
  for (; q + 4 <= qe; q += 4) {
    const double p0 = phi[M.sp_jp[q]], p1 = phi[M.sp_jp[q + 1]];
    const double p2 = phi[M.sp_jp[q + 2]], p3 = phi[M.sp_jp[q + 3]];
    s = s + p0;
    s = s + p1;
    s = s + p2;
    s = s + p3;
  }
  for (; q < qe; ++q) s = s + phi[M.sp_jp[q]];
  return s;
}

// Reaction j at y, direction v: returns dr_j/dy . v and writes, per reactant
// slot q, g * d(dr_j/dy . v)/dy_{react[j,q]} into phi[j*P + q].
__device__ inline double rate_hess(const MechDev& M, const double* y, const double* k, const double* v,
                                   int j, double g, double* phi) {
  const int P = M.P, n = M.n;
  const int* rj = M.react + j * P;
  if (P == 2) {   // straight-line form of the loops below (same operations)
    const int r0 = rj[0], r1 = rj[1];
    const bool h0 = r0 < n, h1 = r1 < n;
    const double kj = k[j];
    const double y0 = h0 ? y[r0] : 1.0, y1 = h1 ? y[r1] : 1.0;
    const double v0 = h0 ? v[r0] : 0.0, v1 = h1 ? v[r1] : 0.0;
    double hk = 0.0;
    if (h0) hk = hk + (kj * y1) * v0;
    if (h1) hk = hk + (kj * y0) * v1;
    phi[j * 2] = g * ((h0 && h1) ? 0.0 + kj * v1 : 0.0);
    phi[j * 2 + 1] = g * ((h0 && h1) ? 0.0 + kj * v0 : 0.0);
    return hk;
  }
  double hk = 0.0;
  for (int p = 0; p < P; ++p)
    if (rj[p] < n) hk = hk + drate(M, y, k, j, p) * v[rj[p]];
  for (int q = 0; q < P; ++q) {
    double acc = 0.0;
    if (rj[q] < n) {
      for (int p = 0; p < P; ++p) {
        if (p == q || rj[p] >= n) continue;
        double prod = k[j];
        for (int t = 0; t < P; ++t)
          if (t != p && t != q && rj[t] < n) prod = prod * y[rj[t]];
        acc = acc + prod * v[rj[p]];
      }
    }
    phi[j * P + q] = g * acc;
  }
  return hk;
}

template <int TILE>
__global__ __launch_bounds__(NT) void k_adjoint(AdjArgs A) {
  extern __shared__ double smem[];
  const MechDev& M = A.M;
  const fode_ctrl_params& p = A.p;
  const fode_adjoint_desc& D = A.a;
  const int n = M.n, R = M.R, NP = M.P, S = p.stages, tid = threadIdx.x;
  const EngineLds L = carve(smem + A.lay.engine, A.P);
  const CellLds V = carve_cell(smem, A.lay, n, R, S);
  double* lam = smem + A.adj;
  double* yb = lam + n;
  double* Kb = yb + n;
  double* Fb = Kb + S * n;
  double* kb = Fb + S * n;
  double* phi = smem + A.phi;
  double* gr = smem + A.gr;
  double* ws = D.ws + (size_t)blockIdx.x * D.ws_stride;
  const long step_len = (long)(S + 1) * n;
  const int cap = (int)(D.ws_stride / step_len);
  for (;;) {
    if (tid == 0) V.flags[0] = atomicAdd(D.queue, 1);
    __syncthreads();
    const int cell = V.flags[0];
    __syncthreads();
    if (cell >= D.C) break;
    const int group = cell / D.group_size;
    const int ns = D.nacc[group];
    const double* hl = D.hlog + (size_t)group * D.hcap;
    const double* kg = D.k + (size_t)cell * D.k_stride;
    if (ns > cap) {   // workspace too small: the cell's results are marked invalid
      for (int i = tid; i < n; i += NT) D.lam[(size_t)cell * n + i] = __builtin_nan("");
      for (int j = tid; j < R; j += NT) D.kbar[(size_t)cell * R + j] = __builtin_nan("");
      __syncthreads();
      continue;
    }
    for (int i = tid; i < n; i += NT) V.y[i] = D.Y[(size_t)cell * n + i];
    for (int j = tid; j < R; j += NT) {
      V.kc[j] = kg[j] * D.kmul[j];
      kb[j] = 0.0;
    }
    __syncthreads();
    // 1. Recompute the accepted steps from the checkpoint; keep each step's
    //    start state and stages.
    for (int s = 0; s < ns; ++s) {
      const double H = hl[s];
      double* w = ws + s * step_len;
      for (int i = tid; i < n; i += NT) w[i] = V.y[i];
      cell_stages<TILE>(A.P, A.P.t, M, p, D.mass, L, V, H, 1.0 / (H * p.gamma), nullptr);
      for (int i = tid; i < S * n; i += NT) w[n + i] = V.K[i];
      for (int i = tid; i < n; i += NT) V.y[i] = V.w[i];
      __syncthreads();
    }
    // 2. The clamp closing the call: d max(y, 0)/dy = [y > 0].
    for (int i = tid; i < n; i += NT) {
      const double l = D.lam[(size_t)cell * n + i];
      lam[i] = V.y[i] > 0.0 ? l : 0.0;
    }
    __syncthreads();
    // 3. The steps, last first.
    for (int s = ns - 1; s >= 0; --s) {
      const double H = hl[s];
      const double alpha = 1.0 / (H * p.gamma);
      const double* w = ws + s * step_len;
      for (int i = tid; i < n; i += NT) V.y[i] = w[i];
      for (int i = tid; i < S * n; i += NT) V.K[i] = w[n + i];
      __syncthreads();
      jacobian_into<slot3_paths(TILE)>(M, V);
      double a[TILE][TILE];
      assemble<TILE>(A.PT, JLds{V.jv}, alpha, D.mass, false, L.lds, a);
      factor_cell<TILE>(A.PT, A.PT.t, L, a);
      for (int i = tid; i < n; i += NT) {
        const double l = lam[i];
        yb[i] = l;
        for (int r = 0; r < S; ++r) {
          Kb[r * n + i] = p.m[r] * l;
          Fb[r * n + i] = 0.0;
        }
      }
      __syncthreads();
      double* vec = L.lds + A.PT.vec;
      double* mu = V.f0;
      for (int r = S - 1; r >= 0; --r) {
        // mu_r = A^{-T} Kbar_r
        for (int q = tid; q < A.PT.c + A.PT.mpad; q += NT) vec[q] = q < n ? Kb[r * n + A.PT.perm[q]] : 0.0;
        solve_cell<TILE>(A.PT, A.PT.t, L, a);
        for (int q = tid; q < n; q += NT) mu[A.PT.perm[q]] = vec[q];
        __syncthreads();
        int src = r;
        while (src > 0 && !p.new_forcing[src]) --src;
        for (int i = tid; i < n; i += NT) {
          const double m = mu[i];
          Fb[src * n + i] += m;
          for (int s2 = 0; s2 < r; ++s2) {
            const double coef = p.c[r * 4 + s2] / H;
            if (coef != 0.0) Kb[s2 * n + i] += coef * m;
          }
        }
        // d/dy [mu^T J(y) K_r] and the rate-constant part of J
        const double* Kr = V.K + r * n;
        for (int j = tid; j < R; j += NT) {
          const double g = st_dot(M, mu, j);
          kb[j] += g * rate_hess(M, V.y, V.kc, Kr, j, g, phi);
        }
        __syncthreads();
        for (int i = tid; i < n; i += NT) yb[i] += gather_slots(M, phi, i);
        __syncthreads();
        if (src == r) {   // a new forcing F(w_r) (stage 0: F(y))
          for (int i = tid; i < n; i += NT) {
            double v = V.y[i];
            for (int s2 = 0; s2 < r; ++s2) {
              const double as = p.a[r * 4 + s2];
              if (as != 0.0) v = v + as * V.K[s2 * n + i];
            }
            V.w[i] = v;
          }
          __syncthreads();
          for (int j = tid; j < R; j += NT) {
            const double g = st_dot(M, Fb + r * n, j);
            gr[j] = g;
            kb[j] += g * rate(M, V.w, V.kc, j);
          }
          __syncthreads();
          for (int jp = tid; jp < R * NP; jp += NT) {
            const int j = jp / NP;
            phi[jp] = M.react[jp] < n ? gr[j] * drate(M, V.w, V.kc, j, jp % NP) : 0.0;
          }
          __syncthreads();
          for (int i = tid; i < n; i += NT) {
            const double wb = gather_slots(M, phi, i);
            yb[i] += wb;
            for (int s2 = 0; s2 < r; ++s2) {
              const double as = p.a[r * 4 + s2];
              if (as != 0.0) Kb[s2 * n + i] += as * wb;
            }
          }
          __syncthreads();
        }
      }
      for (int i = tid; i < n; i += NT) lam[i] = yb[i];
      __syncthreads();
    }
    for (int i = tid; i < n; i += NT) D.lam[(size_t)cell * n + i] = lam[i];
    for (int j = tid; j < R; j += NT) D.kbar[(size_t)cell * R + j] += kb[j];
    __syncthreads();
  }
}

__global__ void k_reset_queue_adj(int* q) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *q = 0;
}

int device_limits(int* cus, int* maxlds) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipDeviceGetAttribute(maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  return 0;
}

int check_pair(const PlanDev& P, const PlanDev& PT, const MechDev& M) {
  if (P.n != M.n || PT.n != M.n || P.nnz != M.nnz || PT.nnz != M.nnz)
    return afail("fode_adjoint: plans and mechanism differ");
  if (P.c != PT.c || P.tile != PT.tile || P.lds_len != PT.lds_len || P.mpad != PT.mpad || P.ns != PT.ns ||
      engine_lds_doubles(P) != engine_lds_doubles(PT))
    return afail("fode_adjoint: the transposed plan's split differs from the plan's");
  if (M.ncon > 0) return afail("fode_adjoint: constrained systems are not supported");
  return 0;
}

template <int TILE>
int blocks_for(const AdjArgs& A, int C, int* blocks, size_t* bytes) {
  int cus = 0, maxlds = 0;
  device_limits(&cus, &maxlds);
  *bytes = (size_t)A.total * sizeof(double);
  if (*bytes > (size_t)maxlds)
    return afail("fode_adjoint: the plan needs " + std::to_string(*bytes) + " bytes of LDS, above " +
                 std::to_string(maxlds));
  hipError_t e = hipFuncSetAttribute((const void*)k_adjoint<TILE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)*bytes);
  if (e != hipSuccess) return afail(std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  int occ = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_adjoint<TILE>, NT, *bytes);
  if (e != hipSuccess) (void)hipGetLastError();   // not the launch's error
  if (e != hipSuccess || occ < 1) occ = 1;
  int b = cus * occ;
  if (b > C) b = C;
  *blocks = b < 1 ? 1 : b;
  return 0;
}

template <int TILE>
int launch_adjoint(AdjArgs A, hipStream_t s) {
  int blocks = 0;
  size_t bytes = 0;
  if (blocks_for<TILE>(A, A.a.C, &blocks, &bytes)) return -1;
  hipLaunchKernelGGL(k_reset_queue_adj, dim3(1), dim3(64), 0, s, A.a.queue);
  hipLaunchKernelGGL(k_adjoint<TILE>, dim3(blocks), dim3(NT), bytes, s, A);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return afail(std::string("k_adjoint: ") + hipGetErrorString(e));
  return 0;
}

}  // namespace

extern "C" {

int fode_adjoint_ws_blocks(void* plan, void* plan_t, void* mech, int stages, int C, int* blocks) {
  if (!plan || !plan_t || !mech || !blocks || stages < 1 || stages > 4 || C < 1)
    return afail("fode_adjoint_ws_blocks: bad argument");
  const PlanDev& P = *static_cast<const PlanDev*>(fode_internal_plan(plan));
  const PlanDev& PT = *static_cast<const PlanDev*>(fode_internal_plan(plan_t));
  const MechDev& M = *static_cast<const MechDev*>(fode_internal_mech(mech));
  if (check_pair(P, PT, M)) return -1;
  const AdjArgs A = make_args(P, PT, M, stages);
  size_t bytes = 0;
  switch (P.tile) {
#define X(T) case T: return blocks_for<T>(A, C, blocks, &bytes);
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10)
#undef X
  }
  return afail("fode_adjoint_ws_blocks: unsupported tile");
}

int fode_adjoint_interval(void* plan, void* plan_t, void* mech, const fode_ctrl_params* p,
                          const fode_adjoint_desc* a, void* stream) {
  if (!plan || !plan_t || !mech || !p || !a) return afail("fode_adjoint_interval: null argument");
  if (a->C <= 0 || a->group_size <= 0 || a->ngroups != (a->C + a->group_size - 1) / a->group_size)
    return afail("fode_adjoint_interval: C, group_size and ngroups disagree");
  if (p->stages < 1 || p->stages > 4) return afail("fode_adjoint_interval: bad tableau");
  if (p->constrained) return afail("fode_adjoint_interval: constrained systems are not supported");
  if (p->loop != FODE_LOOP_ALGORITHM) return afail("fode_adjoint_interval: only Algorithm's step loop is differentiated");
  if (!a->Y || !a->k || !a->kmul || !a->mass || !a->hlog || !a->nacc || !a->lam || !a->kbar || !a->ws ||
      !a->queue)
    return afail("fode_adjoint_interval: null array");
  const PlanDev& P = *static_cast<const PlanDev*>(fode_internal_plan(plan));
  const PlanDev& PT = *static_cast<const PlanDev*>(fode_internal_plan(plan_t));
  const MechDev& M = *static_cast<const MechDev*>(fode_internal_mech(mech));
  if (check_pair(P, PT, M)) return -1;
  if (a->ws_stride < (int64_t)(p->stages + 1) * M.n) return afail("fode_adjoint_interval: ws_stride too small");
  AdjArgs A = make_args(P, PT, M, p->stages);
  A.p = *p;
  A.a = *a;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  switch (P.tile) {
#define X(T) case T: return launch_adjoint<T>(A, s);
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10)
#undef X
  }
  return afail("fode_adjoint_interval: unsupported tile");
}

}  // extern "C"
