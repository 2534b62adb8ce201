// The per-cell LU engine (gfx950, wave64, FP64).
//
// One 256-thread workgroup owns one cell's stage matrix alpha*M - J.  The host
// symbolic plan splits the minimum-degree ordered matrix into
//   * sparse entries (min(i,j) < c), held in LDS and computed by
//     level-scheduled dot-product tasks, and
//   * the dense trailing core (m x m, padded to G*TILE with identity), held in
//     registers, 2-D cyclic over a 16 x 16 thread grid: thread (tr, tc) owns
//     core entries (ti*16 + tr, tj*16 + tc), ti, tj < TILE.
// The core is factored right-looking without pivoting with one barrier per
// two pivots (double-buffered pivot rows/columns in LDS); its 16 x 16 diagonal
// blocks are then inverted into LDS so that each triangular solve is TILE
// block steps of one barrier each.
#pragma once
#include "chem.hpp"
#include "common.hpp"
#include "control.hpp"
#include "timers.hpp"

namespace fode {

// LDS map of the cell engine (doubles), sized by the host from the plan.
struct EngineLds {
  double* lds;    // [lds_len]: sparse entries, then the solve vector (c + mpad)
  double* lbuf;   // [4*mpad]   pivot columns, two per pair, double-buffered
  double* ubuf;   // [4*mpad]   pivot rows
  double* rbuf;   // [mpad]     reciprocals of the first pivot of each pair
  double* pb;     // [2*NT]     block-solve partial products
  double* scr;    // [tile*NT]  diagonal blocks being inverted
  double* dL;     // [tile*NT]  inverse diagonal blocks of L, transposed
  double* dU;     // [tile*NT]  inverse diagonal blocks of U, transposed
};

// After lds: the pivot buffers (lbuf, ubuf, rbuf), then pb, scr, dL, dU, or
// (while the Schur update runs, before any of these is live; the pivot
// buffers are written as soon as a thread leaves it) the Schur sums,
// thread-interleaved S[j * NT + t] from pb on.
__host__ __device__ inline int engine_pb_offset(const PlanDev& P) { return P.lds_len + 9 * P.mpad; }
__host__ __device__ inline int engine_scr_offset(const PlanDev& P) { return engine_pb_offset(P) + 2 * NT; }

__host__ __device__ inline int engine_lds_doubles(const PlanDev& P) {
  const int rest = 2 * NT + 3 * P.tile * NT;
  const int sums = (P.sch_depth + 1) * NT;
  return engine_pb_offset(P) + (rest > sums ? rest : sums);
}

__device__ inline EngineLds carve(double* base, const PlanDev& P) {
  EngineLds e;
  e.lds = base;
  e.lbuf = e.lds + P.lds_len;
  e.ubuf = e.lbuf + 4 * P.mpad;
  e.rbuf = e.ubuf + 4 * P.mpad;
  e.pb = e.rbuf + P.mpad;
  e.scr = e.pb + 2 * NT;
  e.dL = e.scr + P.tile * NT;
  e.dU = e.dL + P.tile * NT;
  return e;
}

// ---------------------------------------------------------------- helpers
// A double moved across lanes by a DPP pattern (both 32-bit halves).
template <int CTRL>
__device__ inline double dpp_mov(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// Sum over the 16 lanes of a DPP row (threads (tr, 0..15) of the engine
// grid), the same bits in every lane: lane ^ 1 and lane ^ 2 by quad
// permutes, then the other quad of the half row (half-row mirror), then the
// other half row (row mirror).  Register-to-register, no LDS crossbar.
__device__ inline double rowsum16(double x) {
  x += dpp_mov<0xB1>(x);    // quad_perm [1, 0, 3, 2]
  x += dpp_mov<0x4E>(x);    // quad_perm [2, 3, 0, 1]
  x += dpp_mov<0x141>(x);   // row_half_mirror
  x += dpp_mov<0x140>(x);   // row_mirror
  return x;
}

// 1/x for pivots (finite, nonzero, normal range): the hardware
// reciprocal refined by two Newton-Raphson steps r += r (1 - x r), the
// residual formed by fma -- a few instructions against the IEEE division's
// scaled sequence, within an ulp of 1/x.
__device__ inline double pivot_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// Level-scheduled dot-product tasks over LDS:
//   lds[t] = (lds[t] - sum_q lds[a_q] * lds[b_q]) [* 1/lds[d]]
// Tasks of one stage are independent; a barrier follows every stage.

// One task per thread: the pairs in order, four pairs' loads in flight at a
// time.
__device__ inline void task_single(const uint2 tk, const uint32_t* pairs, double* lds) {
  const uint32_t tgt = tk.x & 0xFFFFu, div = tk.x >> 16;
  const uint32_t q0 = tk.y & 0xFFFFFu, qe = q0 + (tk.y >> 20);
  double acc = lds[tgt];
  uint32_t q = q0;
  for (; q + 4 <= qe; q += 4) {
    const uint32_t w0 = pairs[q], w1 = pairs[q + 1], w2 = pairs[q + 2], w3 = pairs[q + 3];
    const double x0 = lds[w0 >> 16], y0 = lds[w0 & 0xFFFFu];
    const double x1 = lds[w1 >> 16], y1 = lds[w1 & 0xFFFFu];
    const double x2 = lds[w2 >> 16], y2 = lds[w2 & 0xFFFFu];
    const double x3 = lds[w3 >> 16], y3 = lds[w3 & 0xFFFFu];
    acc = fma(-x0, y0, acc);
    acc = fma(-x1, y1, acc);
    acc = fma(-x2, y2, acc);
    acc = fma(-x3, y3, acc);
  }
  if (q < qe) {
    // the last one to three pairs: their loads in flight together (indices
    // clamped into the task, the extra products discarded), then applied in
    // order
    const uint32_t last = qe - 1;
    const uint32_t w0 = pairs[q], w1 = pairs[min(q + 1, last)], w2 = pairs[min(q + 2, last)];
    const double x0 = lds[w0 >> 16], y0 = lds[w0 & 0xFFFFu];
    const double x1 = lds[w1 >> 16], y1 = lds[w1 & 0xFFFFu];
    const double x2 = lds[w2 >> 16], y2 = lds[w2 & 0xFFFFu];
    acc = fma(-x0, y0, acc);
    if (q + 1 < qe) acc = fma(-x1, y1, acc);
    if (q + 2 < qe) acc = fma(-x2, y2, acc);
  }
  if (div != NO_DIV) acc = acc * pivot_rcp(lds[div]);
  lds[tgt] = acc;
}

// One task per quad (stages of at most NT/4 tasks, which are the long ones
// in the solves): lane j of the quad sums the pairs q0 + j, q0 + j + 4, ...
// in order, the four partial sums are combined by DPP quad permutes (the
// same bits in all four lanes), and lane 0 writes
//   lds[t] = (lds[t] - ((s0 + s1) + (s2 + s3))) [* 1/lds[d]].
// Called with every lane of the workgroup active (the DPP reads them all).
__device__ inline void task_quad(const uint2* tasks, int t, int e, const uint32_t* pairs, double* lds) {
  const int j = threadIdx.x & 3;
  const bool live = t < e;
  double part = 0.0;
  uint32_t tgt = 0, div = NO_DIV;
  if (live) {
    const uint2 tk = tasks[t];
    tgt = tk.x & 0xFFFFu;
    div = tk.x >> 16;
    const uint32_t q0 = tk.y & 0xFFFFFu, qe = q0 + (tk.y >> 20);
    uint32_t q = q0 + j;
    for (; q + 12 < qe; q += 16) {
      const uint32_t w0 = pairs[q], w1 = pairs[q + 4], w2 = pairs[q + 8], w3 = pairs[q + 12];
      const double x0 = lds[w0 >> 16], y0 = lds[w0 & 0xFFFFu];
      const double x1 = lds[w1 >> 16], y1 = lds[w1 & 0xFFFFu];
      const double x2 = lds[w2 >> 16], y2 = lds[w2 & 0xFFFFu];
      const double x3 = lds[w3 >> 16], y3 = lds[w3 & 0xFFFFu];
      part = fma(x0, y0, part);
      part = fma(x1, y1, part);
      part = fma(x2, y2, part);
      part = fma(x3, y3, part);
    }
    if (q < qe) {
      // this lane's last one to three pairs (q, q + 4, q + 8), loads in
      // flight together as in task_single
      const uint32_t last = qe - 1;
      const uint32_t w0 = pairs[q], w1 = pairs[min(q + 4, last)], w2 = pairs[min(q + 8, last)];
      const double x0 = lds[w0 >> 16], y0 = lds[w0 & 0xFFFFu];
      const double x1 = lds[w1 >> 16], y1 = lds[w1 & 0xFFFFu];
      const double x2 = lds[w2 >> 16], y2 = lds[w2 & 0xFFFFu];
      part = fma(x0, y0, part);
      if (q + 4 < qe) part = fma(x1, y1, part);
      if (q + 8 < qe) part = fma(x2, y2, part);
    }
  }
  part += dpp_mov<0xB1>(part);   // quad_perm [1, 0, 3, 2]
  part += dpp_mov<0x4E>(part);   // quad_perm [2, 3, 0, 1]
  if (live && j == 0) {
    double acc = lds[tgt] - part;
    if (div != NO_DIV) acc = acc * pivot_rcp(lds[div]);
    lds[tgt] = acc;
  }
}

// quad: stages of at most NT/4 tasks take a quad per task (uniform choice
// per stage, made from the stage's size alone).
template <bool QUAD>
__device__ inline void run_tasks(const uint2* tasks, const int* ptr, int nst, const uint32_t* pairs,
                                 double* lds) {
  for (int s = 0; s < nst; ++s) {
    const int b = ptr[s], e = ptr[s + 1];
    if (QUAD && 4 * (e - b) <= NT) {
      task_quad(tasks, b + (int)(threadIdx.x >> 2), e, pairs, lds);
    } else {
      for (int t = b + (int)threadIdx.x; t < e; t += NT) task_single(tasks[t], pairs, lds);
    }
    __syncthreads();
  }
}

// Copy n 32-bit words global -> LDS with the whole workgroup.
__device__ inline void copy_words(uint32_t* dst, const uint32_t* src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// Stage the plan's tables into LDS (words at w); returns LDS-pointing tables.
__device__ inline Tables stage_tables(const PlanDev& P, uint32_t* w) {
  Tables t;
  uint32_t* p = w;
  auto take = [&](const void* src, int words) {
    uint32_t* d = p;
    if (words > 0) copy_words(d, static_cast<const uint32_t*>(src), words);
    p += words;
    return d;
  };
  t.f_tasks = reinterpret_cast<const uint2*>(take(P.t.f_tasks, 2 * P.n_f_tasks));
  t.f_ptr = reinterpret_cast<const int*>(take(P.t.f_ptr, P.f_nst + 1));
  t.f_pairs = take(P.t.f_pairs, P.n_f_pairs);
  if ((p - w) & 1) ++p;
  t.sf_tasks = reinterpret_cast<const uint2*>(take(P.t.sf_tasks, 2 * P.n_sf_tasks));
  t.sf_ptr = reinterpret_cast<const int*>(take(P.t.sf_ptr, P.sf_nst + 1));
  if ((p - w) & 1) ++p;
  t.sb_tasks = reinterpret_cast<const uint2*>(take(P.t.sb_tasks, 2 * P.n_sb_tasks));
  t.sb_ptr = reinterpret_cast<const int*>(take(P.t.sb_ptr, P.sb_nst + 1));
  t.s_pairs = take(P.t.s_pairs, P.n_s_pairs);
  p += (4 - ((p - w) & 3)) & 3;   // 16-byte groups of the Schur lists
  t.sch_pairs = take(P.t.sch_pairs, P.n_sch_pairs);
  __syncthreads();
  return t;
}

// Stage only the solves' tables (their task lists, stage bounds and pairs)
// into LDS at w; the factorization's tables stay in memory.  For plans whose
// solves run many small level-scheduled stages, each stage's chain of table
// reads then starts from LDS, where the whole set fits beside two cells.
__device__ inline Tables stage_solve_tables(const PlanDev& P, uint32_t* w) {
  Tables t = P.t;
  uint32_t* p = w;
  auto take = [&](const void* src, int words) {
    uint32_t* d = p;
    if (words > 0) copy_words(d, static_cast<const uint32_t*>(src), words);
    p += words;
    return d;
  };
  t.sf_tasks = reinterpret_cast<const uint2*>(take(P.t.sf_tasks, 2 * P.n_sf_tasks));
  t.sf_ptr = reinterpret_cast<const int*>(take(P.t.sf_ptr, P.sf_nst + 1));
  if ((p - w) & 1) ++p;
  t.sb_tasks = reinterpret_cast<const uint2*>(take(P.t.sb_tasks, 2 * P.n_sb_tasks));
  t.sb_ptr = reinterpret_cast<const int*>(take(P.t.sb_ptr, P.sb_nst + 1));
  t.s_pairs = take(P.t.s_pairs, P.n_s_pairs);
  __syncthreads();
  return t;
}

__host__ __device__ inline int solve_table_words(int n_sf_tasks, int sf_nst, int n_sb_tasks, int sb_nst, int n_s_pairs) {
  int w = 2 * n_sf_tasks + sf_nst + 1;
  w += w & 1;
  return w + 2 * n_sb_tasks + sb_nst + 1 + n_s_pairs;
}

__host__ __device__ inline int table_words(int n_f_tasks, int f_nst, int n_f_pairs, int n_sf_tasks,
                                           int sf_nst, int n_sb_tasks, int sb_nst, int n_s_pairs,
                                           int n_sch_pairs) {
  int w = 2 * n_f_tasks + f_nst + 1 + n_f_pairs;
  w += w & 1;
  w += 2 * n_sf_tasks + sf_nst + 1;
  w += w & 1;
  w += 2 * n_sb_tasks + sb_nst + 1 + n_s_pairs;
  w += (4 - (w & 3)) & 3;
  w += n_sch_pairs;
  return w;
}

// ------------------------------------------------------------- assembly
// Value of entry (J index e) of alpha*diag(mass) - J, formed as
// fastode/linalg.py:24-32 and :60-67 form it: -J, then + alpha*mass on the
// diagonal; alg_only keeps J only on algebraic rows (mass == 0).
template <class JS>
__device__ inline double stage_entry(const JS& jv, int e, bool row_alg, bool diag, double alpha,
                                     double mass_d, bool alg_only) {
  double v = 0.0;
  if (e >= 0) {
    double j = jv(e);
    if (alg_only && !row_alg) j = 0.0;
    v = -j;
  }
  if (diag) v = v + alpha * mass_d;
  return v;
}

// stage_entry without branches, for the unrolled core slots: every read is
// made (entry 0 stands in for a missing one) and the result selected, so
// that the slots' loads issue together.  Same bits as stage_entry.
template <class JS>
__device__ inline double stage_entry_select(const JS& jv, int e, bool row_alg, bool diag, double alpha,
                                            double mass_d, bool alg_only) {
  const double jr = jv(e >= 0 ? e : 0);
  const double j = (alg_only && !row_alg) ? 0.0 : jr;
  const double v = (e >= 0) ? -j : 0.0;
  return diag ? v + alpha * mass_d : v;
}

template <int TILE, class JS>
__device__ inline void assemble(const PlanDev& P, const JS& jv, double alpha, const double* mass,
                                bool alg_only, double* lds, double (&a)[TILE][TILE]) {
  const int tid = threadIdx.x, tr = tid / G, tc = tid % G;
  // four sparse entries per thread per round, their table reads and then
  // their gathers in flight together (entry 0 stands in past the end), for
  // the small cores, whose plans put thousands of entries in the sparse
  // part.  The large cores keep one entry per round: there the batched
  // loads cost registers that the core factor uses (about 1% slower).
  if constexpr (TILE > 5) {
    for (int s = tid; s < P.ns; s += NT) {
      const int d = P.sp_diag[s];
      lds[s] = stage_entry(jv, P.sp_src[s], mass[P.sp_row[s]] == 0.0, d >= 0, alpha,
                           d >= 0 ? mass[d] : 0.0, alg_only);
    }
  } else for (int s0 = tid; s0 < P.ns; s0 += 4 * NT) {
    int d[4], e[4], r[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int s = s0 + u * NT, ss = s < P.ns ? s : 0;
      d[u] = P.sp_diag[ss];
      e[u] = P.sp_src[ss];
      r[u] = P.sp_row[ss];
    }
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double mr = mass[r[u]], md = mass[d[u] >= 0 ? d[u] : 0];
      v[u] = stage_entry_select(jv, e[u], mr == 0.0, d[u] >= 0, alpha, d[u] >= 0 ? md : 0.0, alg_only);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (s0 + u * NT < P.ns) lds[s0 + u * NT] = v[u];
  }
  // the row of each of this thread's core rows: its mass (alg-only mask, diagonal)
  double rmass[TILE];
#pragma unroll
  for (int ti = 0; ti < TILE; ++ti) {
    const int row = P.core_row[ti * G + tr];
    rmass[ti] = row >= 0 ? mass[row] : 1.0;
  }
  const short* src = P.core_src + (size_t)tid * P.slot_stride;
#pragma unroll
  for (int ti = 0; ti < TILE; ++ti)
#pragma unroll
    for (int tj = 0; tj < TILE; ++tj) {
      const int e = src[ti * TILE + tj];
      const bool diag = (ti == tj) && (tr == tc);
      const double v = stage_entry_select(jv, e, rmass[ti] == 0.0, diag, alpha, rmass[ti], alg_only);
      a[ti][tj] = (e == -2) ? 1.0 : v;
    }
}

// Schur update of the core from the sparse pivots, by the owner of each
// entry: S = 0 - sum_p L_ip U_pj over the entry's pairs (final sparse
// entries in LDS).  One flat pass over the thread's own pairs writes its sums
// in slot order to S[j * NT + t] (pb and the inverse scratch, dead until the
// core is factored), then each is added to its register.  A thread reads only what
// it wrote, so no barrier is needed.
template <int TILE>
__device__ inline void schur(const PlanDev& P, const Tables& T, const double* lds, double* S,
                             double (&a)[TILE][TILE]) {
  const int tid = threadIdx.x;
  const int q1 = P.sch_base[tid + 1];
  double acc = 0.0;
  int j = 0;
  // branch-free: the running sum is stored at its slot after every pair
  // (the closing pair's store is the last one there), then the slot index
  // advances and the sum restarts when the pair closes it
  auto step = [&](uint32_t w, double x, double y) {
    acc = fma(-x, y, acc);
    S[j * NT + tid] = acc;
    const int close = (int)(w >> 31);
    j += close;
    acc = close ? 0.0 : acc;
  };
  // groups of 4 pairs (16-byte aligned; the zero padding after the last
  // close only feeds a sum that is never stored)
  for (int q = P.sch_base[tid]; q < q1; q += 4) {
    const uint4 w = *reinterpret_cast<const uint4*>(T.sch_pairs + q);
    const double x0 = lds[(w.x >> 16) & 0x7FFFu], y0 = lds[w.x & 0xFFFFu];
    const double x1 = lds[(w.y >> 16) & 0x7FFFu], y1 = lds[w.y & 0xFFFFu];
    const double x2 = lds[(w.z >> 16) & 0x7FFFu], y2 = lds[w.z & 0xFFFFu];
    const double x3 = lds[(w.w >> 16) & 0x7FFFu], y3 = lds[w.w & 0xFFFFu];
    step(w.x, x0, y0);
    step(w.y, x1, y1);
    step(w.z, x2, y2);
    step(w.w, x3, y3);
  }
  const uint4 mk = *reinterpret_cast<const uint4*>(P.sch_mask + tid * 4);
  const uint32_t mw[4] = {mk.x, mk.y, mk.z, mk.w};
  j = 0;
#pragma unroll
  for (int ti = 0; ti < TILE; ++ti) {
    uint32_t row = 0;
#pragma unroll
    for (int tj = 0; tj < TILE; ++tj) {
      const int s = ti * TILE + tj;
      row |= ((mw[s >> 5] >> (s & 31)) & 1u) << tj;
    }
    // one branch per tile row; a slot's sum index is j plus the row's set
    // bits before it, so the row's loads are independent and issue together
    // (S has sch_depth + 1 rows, so the unmasked reads stay inside it)
    if (row) {
      double sv[TILE];
#pragma unroll
      for (int tj = 0; tj < TILE; ++tj)
        sv[tj] = S[(j + __builtin_popcount(row & ((1u << tj) - 1u))) * NT + tid];
#pragma unroll
      for (int tj = 0; tj < TILE; ++tj) {
        const uint32_t bit = (row >> tj) & 1u;
        const long long raw = __double_as_longlong(sv[tj]);
        a[ti][tj] = a[ti][tj] + __longlong_as_double(raw & -(long long)bit);
      }
      j += __builtin_popcount(row);
    }
  }
}

// Publish the current column and row of pivots k0 = NB*G + kk and k0 + 1
// (updated through pivot k0 - 1) into pair buffer nb, and the reciprocal of
// pivot k0 (final at that point) at rcs[k0].
template <int TILE, int NB>
__device__ inline void publish_pair(const double (&a)[TILE][TILE], int kk, int nb, double* lbuf,
                                    double* ubuf, double* rcs, int tr, int tc) {
  constexpr int MP = G * TILE;
  double* c = lbuf + 2 * nb * MP;
  double* r = ubuf + 2 * nb * MP;
  const int dc = tc - kk, dr = tr - kk;
  if (dc == 0 || dc == 1) {
#pragma unroll
    for (int ti = NB; ti < TILE; ++ti) c[dc * MP + ti * G + tr] = a[ti][NB];
  }
  if (dr == 0 || dr == 1) {
#pragma unroll
    for (int tj = NB; tj < TILE; ++tj) r[dr * MP + tj * G + tc] = a[NB][tj];
    if (dr == 0 && dc == 0) rcs[NB * G + kk] = pivot_rcp(a[NB][NB]);
  }
}

// The 16 pivots of tile block B, two per barrier.  Every thread applies
// pivot k0 to the published column and row of k1 = k0 + 1 itself (the same
// fma the owners would do, so the bits are those of one pivot per barrier),
// then both rank-1 updates in order.  Only tile row/column B straddle the
// pivots, so only there are the multipliers and pivot-row values masked;
// the owners of the pivot columns keep their raw entries (the update leaves
// them alone: their u is 0) and scale them by the pivots' reciprocals once,
// after the block -- the same product l = raw * (1/pivot) the consumers use.
template <int TILE, int B>
__device__ inline void factor_block(double (&a)[TILE][TILE], double* lbuf, double* ubuf,
                                    double* rcs, int tr, int tc) {
  constexpr int MP = G * TILE;
  double rcown = 0.0;
#pragma nounroll
  for (int kk = 0; kk < G; kk += 2) {
    const int k0 = B * G + kk;
    const int buf = (k0 >> 1) & 1;
    const double* c0 = lbuf + 2 * buf * MP;
    const double* c1 = c0 + MP;
    const double* r0 = ubuf + 2 * buf * MP;
    const double* r1 = r0 + MP;
    const double rc0 = rcs[k0];
    const double l10 = c0[B * G + kk + 1] * rc0;   // multiplier of row k1 for pivot k0
    const double u01 = r0[B * G + kk + 1];         // row k0 at column k1
    const double rc1 = pivot_rcp(fma(-l10, u01, r1[B * G + kk + 1]));
    // the reciprocal of this thread's own pivot column, for the scaling below
    rcown = (tc == kk) ? rc0 : rcown;
    rcown = (tc == kk + 1) ? rc1 : rcown;
    double u0[TILE], u1[TILE];
#pragma unroll
    for (int tj = B; tj < TILE; ++tj) {
      const double x0 = r0[tj * G + tc];
      u0[tj] = x0;
      u1[tj] = fma(-l10, x0, r1[tj * G + tc]);
    }
    u0[B] = (tc > kk) ? u0[B] : 0.0;
    u1[B] = (tc > kk + 1) ? u1[B] : 0.0;
#pragma unroll
    for (int ti = B; ti < TILE; ++ti) {
      double l0 = c0[ti * G + tr] * rc0;
      if (ti == B) l0 = (tr > kk) ? l0 : 0.0;
      double l1 = fma(-l0, u01, c1[ti * G + tr]) * rc1;
      if (ti == B) l1 = (tr > kk + 1) ? l1 : 0.0;
#pragma unroll
      for (int tj = B; tj < TILE; ++tj) {
        a[ti][tj] = fma(-l0, u0[tj], a[ti][tj]);
        a[ti][tj] = fma(-l1, u1[tj], a[ti][tj]);
      }
    }
    if (kk + 2 < G) {
      publish_pair<TILE, B>(a, kk + 2, buf ^ 1, lbuf, ubuf, rcs, tr, tc);
    } else if constexpr (B + 1 < TILE) {
      publish_pair<TILE, B + 1>(a, 0, buf ^ 1, lbuf, ubuf, rcs, tr, tc);
    }
    __syncthreads();
  }
  // the L entries of tile column B (below the pivots' diagonal)
  const double rcb = rcown;
#pragma unroll
  for (int ti = B; ti < TILE; ++ti) {
    if (ti == B) {
      if (tr > tc) a[ti][B] = a[ti][B] * rcb;
    } else {
      a[ti][B] = a[ti][B] * rcb;
    }
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
  publish_pair<TILE, 0>(a, 0, 0, L.lbuf, L.ubuf, L.rbuf, tr, tc);
  __syncthreads();
  factor_blocks<TILE, 0>(a, L.lbuf, L.ubuf, L.rbuf, tr, tc);
}

// Inverses of the diagonal 16x16 blocks of L (unit lower) and U, stored
// transposed in LDS: thread (tr, tc) reads entry (tc, tr) of block J at
// J*NT + tid.
template <int TILE>
__device__ inline void diag_inverses(const double (&a)[TILE][TILE], const EngineLds& L) {
  const int tid = threadIdx.x;
  const int task = tid >> 4, col = tid & 15;
  for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
    for (int J = 0; J < TILE; ++J) L.scr[J * NT + tid] = a[J][J];
    __syncthreads();
    double* out = pass == 0 ? L.dL : L.dU;
    if (task < TILE) {
      const double* M = L.scr + task * NT;
      double x[G];
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
          x[r] = v * pivot_rcp(M[r * G + r]);
        }
      }
      // inverse entry (r, col) goes to col*G + r: thread (tr, tc) reads (tc, tr)
#pragma unroll
      for (int r = 0; r < G; ++r) out[task * NT + col * G + r] = x[r];
    }
    __syncthreads();
  }
}

// x[lane of the even row] + x[lane of the odd row] of each pair of DPP rows
// (lanes 16 apart), and of the two wave halves (lanes 32 apart): gfx950's
// permlane swaps give each lane both operands in registers, low one first,
// so both lanes of a pair form the same sum.
__device__ inline double pair_sum16(double x) {
  const unsigned lo = (unsigned)__double2loint(x), hi = (unsigned)__double2hiint(x);
  const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
}
__device__ inline double pair_sum32(double x) {
  const unsigned lo = (unsigned)__double2loint(x), hi = (unsigned)__double2hiint(x);
  const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
}

// Sum over the 16 grid rows of x (the threads of column tc): first over
// the wave's four rows by register permutes, then over the four waves'
// partials through p (one barrier; 4 LDS reads a thread).  The same bits in
// every thread of the column.
__device__ inline double colsum16(double x, double* p) {
  const int tid = threadIdx.x, lane = tid & 63;
  x = pair_sum32(pair_sum16(x));
  if (lane < G) p[(tid >> 6) * G + lane] = x;
  __syncthreads();
  const int tc = tid % G;
  return (p[tc] + p[G + tc]) + (p[2 * G + tc] + p[3 * G + tc]);
}

// Block forward/back substitution of the core: v (length mpad) in LDS holds
// the right-hand side on entry and the solution on exit.
template <int TILE>
__device__ inline void core_solve(const double (&a)[TILE][TILE], const EngineLds& L, double* v) {
  const int tid = threadIdx.x, tr = tid / G, tc = tid % G;
  // each step's inverse entry and right-hand side are read up front (a block
  // of v is written only by its own step, after its reads), so the steps'
  // chains start without an LDS round trip
  double acc[TILE], dv[TILE], vv[TILE];
#pragma unroll
  for (int I = 0; I < TILE; ++I) {
    acc[I] = 0.0;
    dv[I] = L.dL[I * NT + tid];
    vv[I] = v[I * G + tr];
  }
#pragma unroll
  for (int J = 0; J < TILE; ++J) {
    const double d = dv[J];
    const double s = rowsum16(acc[J]);
    const double r = vv[J] - s;
    const double y = colsum16(d * r, L.pb + (J & 1) * NT);
    if (tr == 0) v[J * G + tc] = y;
#pragma unroll
    for (int I = J + 1; I < TILE; ++I) acc[I] = fma(a[I][J], y, acc[I]);
  }
  __syncthreads();
#pragma unroll
  for (int I = 0; I < TILE; ++I) {
    acc[I] = 0.0;
    dv[I] = L.dU[I * NT + tid];
    vv[I] = v[I * G + tr];
  }
#pragma unroll
  for (int J = TILE - 1; J >= 0; --J) {
    const double d = dv[J];
    const double s = rowsum16(acc[J]);
    const double r = vv[J] - s;
    const double x = colsum16(d * r, L.pb + (J & 1) * NT);
    if (tr == 0) v[J * G + tc] = x;
#pragma unroll
    for (int I = 0; I < J; ++I) acc[I] = fma(a[I][J], x, acc[I]);
  }
  __syncthreads();
}

// Full factorization of one cell from assembled values.
template <int TILE>
__device__ __forceinline__ void factor_cell(const PlanDev& P, const Tables& T, const EngineLds& L,
                                   double (&a)[TILE][TILE]) {
  __syncthreads();
  run_tasks<false>(T.f_tasks, T.f_ptr, P.f_nst, T.f_pairs, L.lds);
  FODE_PHASE(PH_SPARSE_FACTOR);
  schur<TILE>(P, T, L.lds, L.pb, a);
  FODE_PHASE(PH_SCHUR);
  core_factor<TILE>(a, L);
  FODE_PHASE(PH_CORE_FACTOR);
  diag_inverses<TILE>(a, L);
  FODE_PHASE(PH_INVERSES);
}

// Solve with the cell's factor; the right-hand side is already in the LDS
// vector (new ordering, padding zero) and the solution is left there.
template <int TILE>
__device__ __forceinline__ void solve_cell(const PlanDev& P, const Tables& T, const EngineLds& L,
                                  const double (&a)[TILE][TILE]) {
  double* vec = L.lds + P.vec;
  __syncthreads();
  FODE_PHASE(PH_STAGE_SETUP);
  run_tasks<true>(T.sf_tasks, T.sf_ptr, P.sf_nst, T.s_pairs, L.lds);
  FODE_PHASE(PH_SPARSE_SOLVE);
  core_solve<TILE>(a, L, vec + P.c);
  FODE_PHASE(PH_CORE_SOLVE);
  run_tasks<true>(T.sb_tasks, T.sb_ptr, P.sb_nst, T.s_pairs, L.lds);
  FODE_PHASE(PH_SPARSE_SOLVE);
}

}  // namespace fode
