// One Rosenbrock attempt on one cell, fused in one workgroup: the cell's
// rates, forcing, Jacobian, stage matrix and its LU, the stage solves and the
// stage combination (fastode/solvers.py:54-86 with the System calls of
// fastode/chem.py:81-134), all in LDS and registers.  Shared by the batched
// persistent kernel and the adjoint's recomputation, so that both produce
// the same bits.
#pragma once
#include "engine.hpp"

namespace fode {

// LDS layout of a workgroup, in doubles from the dynamic base.
struct Layout {
  int tables;      // uint32 words of the staged tables (rounded to doubles)
  int engine;      // engine region offset
  int cell;        // cell region offset
  int rd;          // rates / rate partials: the inverse scratch when it fits
  int jval;        // Jacobian values: the inverse-block region (dead until factored)
  int total;
};

__host__ __device__ inline int cell_region_doubles(int n, int R, int S) {
  return 4 * n + S * n + R + 4 + 2;
}

// Which of the plan's tables a kernel copies into LDS.
enum { STAGE_NONE = 0, STAGE_ALL = 1, STAGE_SOLVE = 2 };

// stage: STAGE_ALL copies every task table into LDS (when they fit beside
// the cell's data), STAGE_SOLVE the solves' tables only.
inline Layout make_layout(const PlanDev& P, const MechDev& M, int S, int stage) {
  Layout l;
  const int words = stage == STAGE_ALL
                        ? table_words(P.n_f_tasks, P.f_nst, P.n_f_pairs, P.n_sf_tasks, P.sf_nst, P.n_sb_tasks,
                                      P.sb_nst, P.n_s_pairs, P.n_sch_pairs)
                        : solve_table_words(P.n_sf_tasks, P.sf_nst, P.n_sb_tasks, P.sb_nst, P.n_s_pairs);
  l.tables = stage != STAGE_NONE ? (words + 1) / 2 : 0;
  l.engine = l.tables;
  int end = l.engine + engine_lds_doubles(P);
  const int scr = P.tile * NT;
  const int scr_at = l.engine + engine_scr_offset(P);
  // Jacobian values are dead once assembled, before the inverse blocks
  // (dL, dU) are written: they start at dL and may run past the engine
  // region.
  l.jval = scr_at + scr;
  if (l.jval + M.nnz > end) end = l.jval + M.nnz;
  // Rates (R) and rate partials (R*P) are live together with the Jacobian
  // values (the partials make them) and between the solves, when only the
  // inverse scratch (scr) is dead; they get a region of their own, after
  // everything else, if scr is too small.
  const int need_rd = M.R * M.P > M.R ? M.R * M.P : M.R;
  if (need_rd <= scr) {
    l.rd = scr_at;
  } else {
    l.rd = end;
    end += need_rd;
  }
  l.cell = end;
  l.total = l.cell + cell_region_doubles(M.n, M.R, S);
  return l;
}

struct CellLds {
  double* y;      // n
  double* f0;     // n   F(Y)
  double* fr;     // n   forcing of a later stage; E at the end
  double* w;      // n   stage point; the new state at the end
  double* K;      // S*n
  double* kc;     // R   rate constants
  double* rd;     // R*P rate partials / R rates
  double* jv;     // nnz Jacobian values
  double* red;    // 4
  int* flags;     // 4 ints
};

__device__ inline CellLds carve_cell(double* smem, const Layout& lay, int n, int R, int S) {
  CellLds V;
  V.rd = smem + lay.rd;
  V.jv = smem + lay.jval;
  V.y = smem + lay.cell;
  V.f0 = V.y + n;
  V.fr = V.f0 + n;
  V.w = V.fr + n;
  V.K = V.w + n;
  V.kc = V.K + S * n;
  V.red = V.kc + R;
  V.flags = reinterpret_cast<int*>(V.red + 4);
  return V;
}

struct JLds {
  const double* J;
  __device__ double operator()(int e) const { return J[e]; }
};

// The cell's rates, forcing and Jacobian, with the same arithmetic (and
// bits) as rate/forcing/drate, but several items per thread per
// round so that their table loads (global) and LDS gathers are in flight
// together instead of one dependent chain after another.  Two-reactant-slot
// mechanisms (P == 2) and three-slot ones (P == 3) take straight-line paths; others the
// generic ones.

// Whether a kernel of core tile TILE takes the three-slot paths.  The large
// cores' kernels leave them out: the code alone shifts their register
// allocation (measured about 0.5% slower on a two-slot mechanism), and the
// generic path gives the same bits.
constexpr bool slot3_paths(int tile) { return tile <= 5; }

template <bool SLOT3 = true>
__device__ inline void rates_into(const MechDev& M, const double* y, const double* kc, double* rates) {
  if (M.P == 2) {
    for (int j0 = threadIdx.x; j0 < M.R; j0 += 2 * NT) {
      double r[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int j = j0 + u * NT, jj = j < M.R ? j : 0;
        const int s0 = M.react[2 * jj], s1 = M.react[2 * jj + 1];
        const double v0 = s0 < M.n ? y[s0] : 1.0, v1 = s1 < M.n ? y[s1] : 1.0;
        r[u] = kc[jj] * (v0 * v1);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u)
        if (j0 + u * NT < M.R) rates[j0 + u * NT] = r[u];
    }
  } else if (SLOT3 && M.P == 3) {
    for (int j0 = threadIdx.x; j0 < M.R; j0 += 2 * NT) {
      double r[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int j = j0 + u * NT, jj = j < M.R ? j : 0;
        const int s0 = M.react[3 * jj], s1 = M.react[3 * jj + 1], s2 = M.react[3 * jj + 2];
        const double v0 = s0 < M.n ? y[s0] : 1.0, v1 = s1 < M.n ? y[s1] : 1.0, v2 = s2 < M.n ? y[s2] : 1.0;
        r[u] = kc[jj] * ((v0 * v1) * v2);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u)
        if (j0 + u * NT < M.R) rates[j0 + u * NT] = r[u];
    }
  } else {
    for (int j = threadIdx.x; j < M.R; j += NT) rates[j] = rate(M, y, kc, j);
  }
  __syncthreads();
}

template <bool WIDE = true>
__device__ inline void forcing_into(const MechDev& M, const double* y, const double* rates, double* out) {
  for (int i = threadIdx.x; i < M.n; i += NT) {
    if (M.con_of_row[i] >= 0) {
      out[i] = forcing(M, y, rates, i);
      continue;
    }
    out[i] = ordered_dot<WIDE>(0.0, rates, M.rhs_rxn, M.rhs_coef, M.rhs_ptr[i], M.rhs_ptr[i + 1]);
  }
  __syncthreads();
}

// Rate partials at V.y into V.rd, then the Jacobian values into V.jv.
template <bool SLOT3 = true>
__device__ inline void jacobian_into(const MechDev& M, const CellLds& V) {
  const int tid = threadIdx.x, RP = M.R * M.P;
  if (M.P == 2) {
    // d rate_j / d y_{react[j,p]} = k_j * y_{react[j,1-p]}
    for (int jp0 = tid; jp0 < RP; jp0 += 4 * NT) {
      double d[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int jp = jp0 + u * NT, jj = jp < RP ? jp : 0;
        const int s = M.react[jj ^ 1];
        d[u] = V.kc[jj >> 1] * (s < M.n ? V.y[s] : 1.0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (jp0 + u * NT < RP) V.rd[jp0 + u * NT] = d[u];
    }
  } else if (SLOT3 && M.P == 3) {
    // d rate_j / d y_{react[j,p]} = (k_j * y_{react[j,a]}) * y_{react[j,b]}, a < b the other slots
    for (int jp0 = tid; jp0 < RP; jp0 += 4 * NT) {
      double d[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int jp = jp0 + u * NT, jj = jp < RP ? jp : 0;
        const int j = jj / 3, p = jj - 3 * j;
        const int sa = M.react[3 * j + (p == 0 ? 1 : 0)], sb = M.react[3 * j + (p == 2 ? 1 : 2)];
        const double va = sa < M.n ? V.y[sa] : 1.0, vb = sb < M.n ? V.y[sb] : 1.0;
        d[u] = (V.kc[j] * va) * vb;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (jp0 + u * NT < RP) V.rd[jp0 + u * NT] = d[u];
    }
  } else {
    for (int jp = tid; jp < RP; jp += NT) V.rd[jp] = drate(M, V.y, V.kc, jp / M.P, jp % M.P);
  }
  __syncthreads();
  // entries: the first term of four entries at once (most entries have one;
  // the term arrays hold at least one element, so index 0 stands in for an
  // entry without terms), then any further terms in order
  for (int e0 = tid; e0 < M.nnz; e0 += 4 * NT) {
    int qa[4], qb[4];
    double v[4], c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = e0 + u * NT, ee = e < M.nnz ? e : 0;
      qa[u] = M.jac_ptr[ee];
      qb[u] = M.jac_ptr[ee + 1];
      c[u] = M.jac_const[ee];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool has = qa[u] < qb[u];
      const int q = has ? qa[u] : 0;
      const double t = M.jac_coef[q] * V.rd[M.jac_src[q]];
      v[u] = has ? 0.0 + t : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (qa[u] + 1 < qb[u]) v[u] = ordered_dot<SLOT3>(v[u], V.rd, M.jac_src, M.jac_coef, qa[u] + 1, qb[u]);
      if (e0 + u * NT < M.nnz) V.jv[e0 + u * NT] = v[u] + c[u];
    }
  }
  __syncthreads();
}

// One attempt of step H (alpha = 1/(H gamma)) from the state in V.y with the
// rate constants in V.kc: stages in V.K, the new state in V.w (and
// ynew_out when given), the embedded error E in V.fr.
template <int TILE>
__device__ __forceinline__ void cell_stages(const PlanDev& P, const Tables& T, const MechDev& M, const fode_ctrl_params& p,
                            const double* mass, const EngineLds& L, const CellLds& V, double H,
                            double alpha, double* ynew_out) {
  const int n = M.n, tid = threadIdx.x;
  // F(Y) (fastode/solvers.py:345) and the Jacobian (:347), recomputed per
  // attempt: they are the same values a retry reuses.
  rates_into<slot3_paths(TILE)>(M, V.y, V.kc, V.rd);
  forcing_into<slot3_paths(TILE)>(M, V.y, V.rd, V.f0);
  FODE_PHASE(PH_RATES);
  jacobian_into<slot3_paths(TILE)>(M, V);
  FODE_PHASE(PH_JACOBIAN);
  double a[TILE][TILE];
  assemble<TILE>(P, JLds{V.jv}, alpha, mass, false, L.lds, a);
  FODE_PHASE(PH_ASSEMBLE);
  factor_cell<TILE>(P, T, L, a);
  // Stages (fastode/solvers.py:54-86).
  double* vec = L.lds + P.vec;
  const int S = p.stages;
  const bool dae = p.constrained != 0;
  const double* forcing_src = V.f0;
  for (int r = 0; r < S; ++r) {
    if (r > 0 && p.new_forcing[r]) {
      for (int i = tid; i < n; i += NT) {
        double v = V.y[i];
        for (int s = 0; s < r; ++s) {
          const double as = p.a[r * 4 + s];
          if (as != 0.0) v = v + as * V.K[s * n + i];
        }
        V.w[i] = v;
      }
      __syncthreads();
      rates_into<slot3_paths(TILE)>(M, V.w, V.kc, V.rd);
      forcing_into<slot3_paths(TILE)>(M, V.w, V.rd, V.fr);
      forcing_src = V.fr;
    }
    // right-hand side of stage r, gathered into the solve vector (new order)
    for (int q = tid; q < P.c + P.mpad; q += NT) {
      double v = 0.0;
      if (q < n) {
        const int i = P.perm[q];
        v = forcing_src[i];
        for (int s = 0; s < r; ++s) {
          const double coef = p.c[r * 4 + s] / H;
          if (coef != 0.0) {
            const double ks = dae ? V.K[s * n + i] * mass[i] : V.K[s * n + i];
            v = v + coef * ks;
          }
        }
      }
      vec[q] = v;
    }
    solve_cell<TILE>(P, T, L, a);
    for (int q = tid; q < n; q += NT) V.K[r * n + P.perm[q]] = vec[q];
    __syncthreads();
    FODE_PHASE(PH_STAGE_SETUP);
  }
  // Ynew and the embedded error; E kept in fr.
  for (int i = tid; i < n; i += NT) {
    double y = V.y[i], e = 0.0;
    for (int r = 0; r < S; ++r) {
      const double kr = V.K[r * n + i];
      if (p.m[r] != 0.0) y = y + p.m[r] * kr;
      if (p.e[r] != 0.0) e = e + p.e[r] * kr;
    }
    V.w[i] = y;
    V.fr[i] = e;
    if (ynew_out) ynew_out[i] = y;
  }
  __syncthreads();
}

}  // namespace fode
