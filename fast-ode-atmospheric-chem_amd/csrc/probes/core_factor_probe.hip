// Measurement probe of the register-tiled core factor (engine.hpp
// core_factor), standalone: not part of the library.
//
//   python fast-ode-atmospheric-chem_amd/build.py --probes
//   fast-ode-atmospheric-chem_amd/csrc/probes/core_factor_probe [reps]
//
// One 256-thread workgroup per CU (the LDS request admits one) factors a
// diagonally dominant core of 16*TILE rows `reps` times; the shader clock
// (s_memtime) brackets each factorization.  Reported per TILE (9: the
// 200-species core, 5: the constrained 100-species core, 2: TS1):
//   * the library's factor (variant 0) and six restructurings of it that
//     keep the fma sequence of every entry (each is checked bit for bit
//     against variant 0 on the same matrices):
//       1 deferred: the pivot block's cross updated pair by pair, the
//         trailing tiles by one rank-16 pass per block;
//       2 lookahead: the next pair published before the bulk update and
//         signalled through an LDS counter instead of a barrier, prefetched
//         under the bulk update;
//       3 publish-first: the barrier scheme with the next pair's values
//         published before the bulk update;
//       4 blocked: the 16x16 diagonal block factored inside each wave (DPP
//         row broadcasts, permlane swaps), the column panel by in-row
//         substitution, the row panel through one LDS transpose, one
//         rank-16 trailing pass; three barriers per block;
//       5 quad: four pivots per barrier instead of two, every thread
//         forming the 4x4 pivot block's multipliers itself;
//       6 dual: two cells' cores per workgroup, factored with the library's
//         pair scheme and sharing each pair's barrier (time per core);
//   * the library's factor with two, three and four workgroups per CU
//     (throughput per factorization);
//   * latency floors of one pivot-pair round: a barrier, an LDS
//     write-barrier-read round trip, that plus the pivot reciprocal, a chain
//     of dependent fp64 fma, the reciprocal chain.
// Every launch's indexing stays inside its buffers (LDS regions sized below
// the request; outputs one tile per thread per workgroup), and the only
// wait besides barriers (variant 2) is bounded.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../engine.hpp"

using namespace fode;

namespace probe {

constexpr int PS = 18;   // panel row stride (doubles): distinct banks for a wave's 16 columns

// ------------------------------------------------------------ 1. deferred
template <int TILE, int B>
__device__ inline void deferred_block(double (&a)[TILE][TILE], double* lbuf, double* ubuf, double* rcs,
                                      double* Lp, double* Up, int tr, int tc) {
  constexpr int MP = G * TILE;
  double rcown = 0.0;
#pragma nounroll
  for (int kk = 0; kk < G; kk += 2) {
    const int k0 = B * G + kk, buf = (k0 >> 1) & 1;
    const double* c0 = lbuf + 2 * buf * MP;
    const double* c1 = c0 + MP;
    const double* r0 = ubuf + 2 * buf * MP;
    const double* r1 = r0 + MP;
    const double rc0 = rcs[k0];
    const double l10 = c0[B * G + kk + 1] * rc0, u01 = r0[B * G + kk + 1];
    const double rc1 = pivot_rcp(fma(-l10, u01, r1[B * G + kk + 1]));
    rcown = (tc == kk) ? rc0 : ((tc == kk + 1) ? rc1 : rcown);
    double u0[TILE], u1[TILE];
#pragma unroll
    for (int tj = B; tj < TILE; ++tj) {
      u0[tj] = r0[tj * G + tc];
      u1[tj] = fma(-l10, u0[tj], r1[tj * G + tc]);
    }
    u0[B] = (tc > kk) ? u0[B] : 0.0;
    u1[B] = (tc > kk + 1) ? u1[B] : 0.0;
    double ls0 = 0.0, ls1 = 0.0;
#pragma unroll
    for (int ti = B; ti < TILE; ++ti) {
      double l0 = c0[ti * G + tr] * rc0;
      if (ti == B) l0 = (tr > kk) ? l0 : 0.0;
      double l1 = fma(-l0, u01, c1[ti * G + tr]) * rc1;
      if (ti == B) l1 = (tr > kk + 1) ? l1 : 0.0;
      if (ti == B) {
#pragma unroll
        for (int tj = B; tj < TILE; ++tj) a[B][tj] = fma(-l1, u1[tj], fma(-l0, u0[tj], a[B][tj]));
      } else {
        a[ti][B] = fma(-l1, u1[B], fma(-l0, u0[B], a[ti][B]));
        ls0 = (tc == ti - B - 1) ? l0 : ls0;
        ls1 = (tc == ti - B - 1) ? l1 : ls1;
      }
    }
    if constexpr (B + 1 < TILE) {
      if (tc < TILE - B - 1) *reinterpret_cast<double2*>(Lp + ((B + 1 + tc) * G + tr) * PS + kk) = make_double2(ls0, ls1);
      double us0 = 0.0, us1 = 0.0;
#pragma unroll
      for (int tj = B + 1; tj < TILE; ++tj) {
        us0 = (tr == tj - B - 1) ? u0[tj] : us0;
        us1 = (tr == tj - B - 1) ? u1[tj] : us1;
      }
      if (tr < TILE - B - 1) *reinterpret_cast<double2*>(Up + ((B + 1 + tr) * G + tc) * PS + kk) = make_double2(us0, us1);
    }
    if (kk + 2 < G) publish_pair<TILE, B>(a, kk + 2, buf ^ 1, lbuf, ubuf, rcs, tr, tc);
    __syncthreads();
  }
#pragma unroll
  for (int ti = B; ti < TILE; ++ti)
    if (ti != B || tr > tc) a[ti][B] = a[ti][B] * rcown;
  if constexpr (B + 1 < TILE) {
#pragma unroll 2
    for (int kq = 0; kq < G; kq += 2) {
      double2 lv[TILE], uv[TILE];
#pragma unroll
      for (int ti = B + 1; ti < TILE; ++ti) lv[ti] = *reinterpret_cast<const double2*>(Lp + (ti * G + tr) * PS + kq);
#pragma unroll
      for (int tj = B + 1; tj < TILE; ++tj) uv[tj] = *reinterpret_cast<const double2*>(Up + (tj * G + tc) * PS + kq);
#pragma unroll
      for (int ti = B + 1; ti < TILE; ++ti)
#pragma unroll
        for (int tj = B + 1; tj < TILE; ++tj)
          a[ti][tj] = fma(-lv[ti].y, uv[tj].y, fma(-lv[ti].x, uv[tj].x, a[ti][tj]));
    }
    publish_pair<TILE, B + 1>(a, 0, 0, lbuf, ubuf, rcs, tr, tc);
    __syncthreads();
  }
}

// --------------------------------------------- 2. lookahead, 3. publish-first
template <int TILE>
struct PairRaw {
  double c0[TILE], c1[TILE], r0[TILE], r1[TILE];
  double rc0, qc0, qr0, qr1;
};

__device__ unsigned long long g_spin_timeouts;

// Wait until the LDS counter reaches target; bounded (a timeout is counted
// and releases every later wait of the workgroup).
__device__ inline void la_wait(int* cnt, int target) {
  int spins = 0;
  while (__hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < target) {
    __builtin_amdgcn_s_sleep(1);
    if (++spins > (1 << 16)) {
      if ((threadIdx.x & 63) == 0) {
        atomicAdd(&g_spin_timeouts, 1ull);
        __hip_atomic_fetch_add(cnt, 1 << 24, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      break;
    }
  }
}

template <int TILE, int B>
__device__ inline void la_load(PairRaw<TILE>& w, const double* lbuf, const double* ubuf, const double* rcs, int slot,
                               int kk, int tr, int tc) {
  constexpr int MP = G * TILE;
  const double* c0 = lbuf + 2 * slot * MP;
  const double* r0 = ubuf + 2 * slot * MP;
  w.rc0 = rcs[B * G + kk];
  w.qc0 = c0[B * G + kk + 1];
  w.qr0 = r0[B * G + kk + 1];
  w.qr1 = r0[MP + B * G + kk + 1];
#pragma unroll
  for (int ti = B; ti < TILE; ++ti) {
    w.c0[ti] = c0[ti * G + tr];
    w.c1[ti] = c0[MP + ti * G + tr];
  }
#pragma unroll
  for (int tj = B; tj < TILE; ++tj) {
    w.r0[tj] = r0[tj * G + tc];
    w.r1[tj] = r0[MP + tj * G + tc];
  }
}

// The pair (NB, kk2) into slot s: its values after the current pair's
// update (multipliers l0/l1 of this thread's rows, u0/u1 of its columns).
template <int TILE, int NB>
__device__ inline void la_publish(const double (&a)[TILE][TILE], const double (&l0)[TILE], const double (&l1)[TILE],
                                  const double (&u0)[TILE], const double (&u1)[TILE], int kk2, int s, double* lbuf,
                                  double* ubuf, double* rcs, int tr, int tc) {
  constexpr int MP = G * TILE;
  double* c = lbuf + 2 * s * MP;
  double* r = ubuf + 2 * s * MP;
  const int dc = tc - kk2, dr = tr - kk2;
  if (dc == 0 || dc == 1) {
#pragma unroll
    for (int ti = NB; ti < TILE; ++ti) c[dc * MP + ti * G + tr] = fma(-l1[ti], u1[NB], fma(-l0[ti], u0[NB], a[ti][NB]));
  }
  if (dr == 0 || dr == 1) {
#pragma unroll
    for (int tj = NB; tj < TILE; ++tj) r[dr * MP + tj * G + tc] = fma(-l1[NB], u1[tj], fma(-l0[NB], u0[tj], a[NB][tj]));
    if (dr == 0 && dc == 0) rcs[NB * G + kk2] = pivot_rcp(fma(-l1[NB], u1[NB], fma(-l0[NB], u0[NB], a[NB][NB])));
  }
}

// One pair's multipliers from its published values (the library's sequence).
template <int TILE, int B>
__device__ inline void pair_multipliers(const PairRaw<TILE>& w, int kk, int tr, int tc, double (&l0)[TILE],
                                        double (&l1)[TILE], double (&u0)[TILE], double (&u1)[TILE], double& rcown) {
  const double rc0 = w.rc0, l10 = w.qc0 * rc0, u01 = w.qr0;
  const double rc1 = pivot_rcp(fma(-l10, u01, w.qr1));
  rcown = (tc == kk) ? rc0 : ((tc == kk + 1) ? rc1 : rcown);
#pragma unroll
  for (int tj = B; tj < TILE; ++tj) {
    u0[tj] = w.r0[tj];
    u1[tj] = fma(-l10, w.r0[tj], w.r1[tj]);
  }
  u0[B] = (tc > kk) ? u0[B] : 0.0;
  u1[B] = (tc > kk + 1) ? u1[B] : 0.0;
#pragma unroll
  for (int ti = B; ti < TILE; ++ti) {
    double x0 = w.c0[ti] * rc0;
    if (ti == B) x0 = (tr > kk) ? x0 : 0.0;
    double x1 = fma(-x0, u01, w.c1[ti]) * rc1;
    if (ti == B) x1 = (tr > kk + 1) ? x1 : 0.0;
    l0[ti] = x0;
    l1[ti] = x1;
  }
}

template <int TILE, int B>
__device__ inline void pair_update(double (&a)[TILE][TILE], const double (&l0)[TILE], const double (&l1)[TILE],
                                   const double (&u0)[TILE], const double (&u1)[TILE]) {
#pragma unroll
  for (int ti = B; ti < TILE; ++ti)
#pragma unroll
    for (int tj = B; tj < TILE; ++tj) a[ti][tj] = fma(-l1[ti], u1[tj], fma(-l0[ti], u0[tj], a[ti][tj]));
}

template <int TILE, int B>
__device__ inline void scale_column(double (&a)[TILE][TILE], double rcown, int tr, int tc) {
#pragma unroll
  for (int ti = B; ti < TILE; ++ti)
    if (ti != B || tr > tc) a[ti][B] = a[ti][B] * rcown;
}

// LOOKAHEAD: counters (four pair slots) instead of barriers; else the
// barrier scheme with the next pair published first.
template <int TILE, int B, bool LOOKAHEAD>
__device__ inline void ahead_block(double (&a)[TILE][TILE], PairRaw<TILE>& w, double* lbuf, double* ubuf, double* rcs,
                                   int* cnt, int tr, int tc) {
  double rcown = 0.0;
#pragma nounroll
  for (int kk = 0; kk < G; kk += 2) {
    const int p = (B * G + kk) >> 1;
    if constexpr (!LOOKAHEAD) la_load<TILE, B>(w, lbuf, ubuf, rcs, p & 1, kk, tr, tc);
    double u0[TILE], u1[TILE], l0[TILE], l1[TILE];
    pair_multipliers<TILE, B>(w, kk, tr, tc, l0, l1, u0, u1, rcown);
    const int slot = LOOKAHEAD ? ((p + 1) & 3) : ((p + 1) & 1);
    if (kk + 2 < G) {
      la_publish<TILE, B>(a, l0, l1, u0, u1, kk + 2, slot, lbuf, ubuf, rcs, tr, tc);
      if constexpr (LOOKAHEAD) {
        if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        la_wait(cnt, 4 * (p + 1));
        la_load<TILE, B>(w, lbuf, ubuf, rcs, slot, kk + 2, tr, tc);
      }
    } else if constexpr (B + 1 < TILE) {
      la_publish<TILE, B + 1>(a, l0, l1, u0, u1, 0, slot, lbuf, ubuf, rcs, tr, tc);
      if constexpr (LOOKAHEAD) {
        if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        la_wait(cnt, 4 * (p + 1));
        la_load<TILE, B + 1>(w, lbuf, ubuf, rcs, slot, 0, tr, tc);
      }
    }
    pair_update<TILE, B>(a, l0, l1, u0, u1);
    if constexpr (!LOOKAHEAD) __syncthreads();
  }
  scale_column<TILE, B>(a, rcown, tr, tc);
}

// -------------------------------------------------------------- 4. blocked
template <int K>
__device__ inline double row_bcast(double x) {   // lane K of each 16-lane row
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x150 + K, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x150 + K, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ inline double read_lane(double x, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), lane),
                          __builtin_amdgcn_readlane(__double2loint(x), lane));
}

template <int R>
__device__ inline double group_bcast(double x) {   // row group R's value, same column lane
  const unsigned lo = (unsigned)__double2loint(x), hi = (unsigned)__double2hiint(x);
  const auto l16 = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto h16 = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  const unsigned al = (R & 1) ? l16[1] : l16[0], ah = (R & 1) ? h16[1] : h16[0];
  const auto l32 = __builtin_amdgcn_permlane32_swap(al, al, false, false);
  const auto h32 = __builtin_amdgcn_permlane32_swap(ah, ah, false, false);
  return __hiloint2double((int)((R & 2) ? h32[1] : h32[0]), (int)((R & 2) ? l32[1] : l32[0]));
}

// Pivot K of the diagonal block held as d[i] = (row 4q+i, column j).
template <int K>
__device__ inline void diag_pivots(double (&d)[4], int q, int j, double (&rc)[G]) {
  if constexpr (K < G) {
    const double r = pivot_rcp(read_lane(d[K % 4], 16 * (K / 4) + K));
    rc[K] = r;
    const double u = group_bcast<K / 4>(d[K % 4]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double t = d[i] * r;
      const double l = row_bcast<K>(t);
      const double upd = fma(-l, u, d[i]);
      d[i] = (4 * q + i > K) ? (j == K ? t : (j > K ? upd : d[i])) : d[i];
    }
    diag_pivots<K + 1>(d, q, j, rc);
  }
}

template <int TILE, int B, int P>
__device__ inline void column_steps(double (&a)[TILE][TILE], const double (&ucol)[G], const double (&rc)[G], int tc) {
  if constexpr (P < G) {
#pragma unroll
    for (int ti = B + 1; ti < TILE; ++ti) {
      const double l = row_bcast<P>(a[ti][B]) * rc[P];
      const double upd = fma(-l, ucol[P], a[ti][B]);
      a[ti][B] = tc > P ? upd : (tc == P ? l : a[ti][B]);
    }
    column_steps<TILE, B, P + 1>(a, ucol, rc, tc);
  }
}

// D: the gathered diagonal block; F: its factors (every wave writes the
// same bits); Lp, Up: the panels.
template <int TILE, int B>
__device__ inline void blocked_block(double (&a)[TILE][TILE], double* D, double* F, double* Lp, double* Up, int tr,
                                     int tc) {
  const int tid = threadIdx.x, lane = tid & 63, q = lane >> 4, j = lane & 15;
  double d[4], rc[G];
#pragma unroll
  for (int i = 0; i < 4; ++i) d[i] = D[(4 * q + i) * G + j];
  diag_pivots<0>(d, q, j, rc);
#pragma unroll
  for (int i = 0; i < 4; ++i) F[(4 * q + i) * G + j] = d[i];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  a[B][B] = F[tr * G + tc];
  if constexpr (B + 1 < TILE) {
    double ucol[G];
#pragma unroll
    for (int p = 0; p < G; ++p) ucol[p] = F[p * G + tc];
    column_steps<TILE, B, 0>(a, ucol, rc, tc);
#pragma unroll
    for (int ti = B + 1; ti < TILE; ++ti) Lp[(ti * G + tr) * PS + tc] = a[ti][B];
#pragma unroll
    for (int tj = B + 1; tj < TILE; ++tj) Up[(tj * G + tc) * PS + tr] = a[B][tj];
    __syncthreads();
    constexpr int NC = (TILE - B - 1) * G;
    if (tid < NC) {
      double* col = Up + ((B + 1) * G + tid) * PS;
      double x[G];
#pragma unroll
      for (int k = 0; k < G; ++k) x[k] = col[k];
#pragma unroll
      for (int p = 0; p < G - 1; ++p)
#pragma unroll
        for (int k = p + 1; k < G; ++k) x[k] = fma(-F[k * G + p], x[p], x[k]);
#pragma unroll
      for (int k = 0; k < G; ++k) col[k] = x[k];
    }
    __syncthreads();
#pragma unroll
    for (int tj = B + 1; tj < TILE; ++tj) a[B][tj] = Up[(tj * G + tc) * PS + tr];
#pragma unroll 2
    for (int kq = 0; kq < G; kq += 2) {
      double2 lv[TILE], uv[TILE];
#pragma unroll
      for (int ti = B + 1; ti < TILE; ++ti) lv[ti] = *reinterpret_cast<const double2*>(Lp + (ti * G + tr) * PS + kq);
#pragma unroll
      for (int tj = B + 1; tj < TILE; ++tj) uv[tj] = *reinterpret_cast<const double2*>(Up + (tj * G + tc) * PS + kq);
#pragma unroll
      for (int ti = B + 1; ti < TILE; ++ti)
#pragma unroll
        for (int tj = B + 1; tj < TILE; ++tj)
          a[ti][tj] = fma(-lv[ti].y, uv[tj].y, fma(-lv[ti].x, uv[tj].x, a[ti][tj]));
    }
    D[tid] = a[B + 1][B + 1];
  }
  __syncthreads();
}


// ------------------------------------------------------------------ 5. quad
// Four pivots per barrier: the columns and rows of pivots k0..k0+3 (updated
// through k0-1) are published together, and every thread forms the 4x4
// pivot block's multipliers and reciprocals itself, in the order the pair
// scheme forms them.
template <int TILE, int NB>
__device__ inline void publish_quad(const double (&a)[TILE][TILE], int kk, int nb, double* lbuf, double* ubuf,
                                    double* rcs, int tr, int tc) {
  constexpr int MP = G * TILE;
  double* c = lbuf + 4 * nb * MP;
  double* r = ubuf + 4 * nb * MP;
  const int dc = tc - kk, dr = tr - kk;
  if (dc >= 0 && dc < 4) {
#pragma unroll
    for (int ti = NB; ti < TILE; ++ti) c[dc * MP + ti * G + tr] = a[ti][NB];
  }
  if (dr >= 0 && dr < 4) {
#pragma unroll
    for (int tj = NB; tj < TILE; ++tj) r[dr * MP + tj * G + tc] = a[NB][tj];
    if (dr == 0 && dc == 0) rcs[NB * G + kk] = pivot_rcp(a[NB][NB]);
  }
}

template <int TILE, int B>
__device__ inline void quad_block(double (&a)[TILE][TILE], double* lbuf, double* ubuf, double* rcs, int tr, int tc) {
  constexpr int MP = G * TILE;
  double rcown = 0.0;
#pragma nounroll
  for (int kk = 0; kk < G; kk += 4) {
    const int k0 = B * G + kk, buf = (k0 >> 2) & 1;
    const double* c0 = lbuf + 4 * buf * MP;
    const double* c1 = c0 + MP;
    const double* c2 = c1 + MP;
    const double* c3 = c2 + MP;
    const double* r0 = ubuf + 4 * buf * MP;
    const double* r1 = r0 + MP;
    const double* r2 = r1 + MP;
    const double* r3 = r2 + MP;
    const int q = B * G + kk;
    // the 4x4 pivot block
    const double rc0 = rcs[k0];
    const double l10 = c0[q + 1] * rc0, l20 = c0[q + 2] * rc0, l30 = c0[q + 3] * rc0;
    const double u01 = r0[q + 1], u02 = r0[q + 2], u03 = r0[q + 3];
    const double rc1 = pivot_rcp(fma(-l10, u01, r1[q + 1]));
    const double u12 = fma(-l10, u02, r1[q + 2]), u13 = fma(-l10, u03, r1[q + 3]);
    const double l21 = fma(-l20, u01, c1[q + 2]) * rc1, l31 = fma(-l30, u01, c1[q + 3]) * rc1;
    const double rc2 = pivot_rcp(fma(-l21, u12, fma(-l20, u02, r2[q + 2])));
    const double u23 = fma(-l21, u13, fma(-l20, u03, r2[q + 3]));
    const double l32 = fma(-l31, u12, fma(-l30, u02, c2[q + 3])) * rc2;
    const double rc3 = pivot_rcp(fma(-l32, u23, fma(-l31, u13, fma(-l30, u03, r3[q + 3]))));
    rcown = (tc == kk) ? rc0 : ((tc == kk + 1) ? rc1 : ((tc == kk + 2) ? rc2 : ((tc == kk + 3) ? rc3 : rcown)));
    double v0[TILE], v1[TILE], v2[TILE], v3[TILE];
#pragma unroll
    for (int tj = B; tj < TILE; ++tj) {
      const double w0 = r0[tj * G + tc];
      const double w1 = fma(-l10, w0, r1[tj * G + tc]);
      const double w2 = fma(-l21, w1, fma(-l20, w0, r2[tj * G + tc]));
      const double w3 = fma(-l32, w2, fma(-l31, w1, fma(-l30, w0, r3[tj * G + tc])));
      v0[tj] = w0;
      v1[tj] = w1;
      v2[tj] = w2;
      v3[tj] = w3;
    }
    v0[B] = (tc > kk) ? v0[B] : 0.0;
    v1[B] = (tc > kk + 1) ? v1[B] : 0.0;
    v2[B] = (tc > kk + 2) ? v2[B] : 0.0;
    v3[B] = (tc > kk + 3) ? v3[B] : 0.0;
#pragma unroll
    for (int ti = B; ti < TILE; ++ti) {
      double x0 = c0[ti * G + tr] * rc0;
      if (ti == B) x0 = (tr > kk) ? x0 : 0.0;
      double x1 = fma(-x0, u01, c1[ti * G + tr]) * rc1;
      if (ti == B) x1 = (tr > kk + 1) ? x1 : 0.0;
      double x2 = fma(-x1, u12, fma(-x0, u02, c2[ti * G + tr])) * rc2;
      if (ti == B) x2 = (tr > kk + 2) ? x2 : 0.0;
      double x3 = fma(-x2, u23, fma(-x1, u13, fma(-x0, u03, c3[ti * G + tr]))) * rc3;
      if (ti == B) x3 = (tr > kk + 3) ? x3 : 0.0;
#pragma unroll
      for (int tj = B; tj < TILE; ++tj) {
        double t = fma(-x0, v0[tj], a[ti][tj]);
        t = fma(-x1, v1[tj], t);
        t = fma(-x2, v2[tj], t);
        a[ti][tj] = fma(-x3, v3[tj], t);
      }
    }
    if (kk + 4 < G) {
      publish_quad<TILE, B>(a, kk + 4, buf ^ 1, lbuf, ubuf, rcs, tr, tc);
    } else if constexpr (B + 1 < TILE) {
      publish_quad<TILE, B + 1>(a, 0, buf ^ 1, lbuf, ubuf, rcs, tr, tc);
    }
    __syncthreads();
  }
#pragma unroll
  for (int ti = B; ti < TILE; ++ti)
    if (ti != B || tr > tc) a[ti][B] = a[ti][B] * rcown;
}

template <int TILE, int V, int B>
__device__ inline void variant_blocks(double (&a)[TILE][TILE], double* smem, PairRaw<TILE>& w, int tr, int tc) {
  if constexpr (B < TILE) {
    constexpr int MP = G * TILE;
    if constexpr (V == 1)
      deferred_block<TILE, B>(a, smem, smem + 4 * MP, smem + 8 * MP, smem + 9 * MP, smem + 9 * MP + MP * PS, tr, tc);
    if constexpr (V == 2)
      ahead_block<TILE, B, true>(a, w, smem, smem + 8 * MP, smem + 16 * MP, reinterpret_cast<int*>(smem + 17 * MP),
                                 tr, tc);
    if constexpr (V == 3) ahead_block<TILE, B, false>(a, w, smem, smem + 4 * MP, smem + 8 * MP, nullptr, tr, tc);
    if constexpr (V == 4) blocked_block<TILE, B>(a, smem, smem + NT, smem + 2 * NT, smem + 2 * NT + MP * PS, tr, tc);
    if constexpr (V == 5) quad_block<TILE, B>(a, smem, smem + 8 * MP, smem + 16 * MP, tr, tc);
    variant_blocks<TILE, V, B + 1>(a, smem, w, tr, tc);
  }
}

// LDS (doubles): V1 9*MP + 2*MP*PS; V2 17*MP + 1; V3 9*MP; V4 2*NT + 2*MP*PS;
// V5 17*MP; V6 (two cores) 18*MP: at most 6,624 at TILE 10, far below the request.
template <int TILE, int V>
__device__ inline void factor_variant(double (&a)[TILE][TILE], double* smem) {
  const int tid = threadIdx.x, tr = tid / G, tc = tid % G;
  constexpr int MP = G * TILE;
  PairRaw<TILE> w;
  if constexpr (V == 4) {
    smem[tid] = a[0][0];
    __syncthreads();
  } else if constexpr (V == 5) {
    publish_quad<TILE, 0>(a, 0, 0, smem, smem + 8 * MP, smem + 16 * MP, tr, tc);
    __syncthreads();
  } else {
    double* rcs = smem + (V == 2 ? 16 : 8) * MP;
    publish_pair<TILE, 0>(a, 0, 0, smem, smem + (V == 2 ? 8 : 4) * MP, rcs, tr, tc);
    if constexpr (V == 2)
      if (tid == 0) *reinterpret_cast<int*>(smem + 17 * MP) = 0;
    __syncthreads();
    if constexpr (V == 2) la_load<TILE, 0>(w, smem, smem + 8 * MP, rcs, 0, 0, tr, tc);
  }
  variant_blocks<TILE, V, 0>(a, smem, w, tr, tc);
}


// ------------------------------------------------------------------ 6. dual
// Two cells' cores factored in one workgroup, each with the library's pair
// scheme and buffers of its own, their pair rounds sharing one barrier so
// that one core's chain overlaps the other's.
template <int TILE, int B>
__device__ inline void pair_round(double (&a)[TILE][TILE], int kk, double* lbuf, double* ubuf, double* rcs, int tr,
                                  int tc, double& rcown) {
  constexpr int MP = G * TILE;
  const int k0 = B * G + kk, buf = (k0 >> 1) & 1;
  const double* c0 = lbuf + 2 * buf * MP;
  const double* c1 = c0 + MP;
  const double* r0 = ubuf + 2 * buf * MP;
  const double* r1 = r0 + MP;
  const double rc0 = rcs[k0];
  const double l10 = c0[B * G + kk + 1] * rc0, u01 = r0[B * G + kk + 1];
  const double rc1 = pivot_rcp(fma(-l10, u01, r1[B * G + kk + 1]));
  rcown = (tc == kk) ? rc0 : ((tc == kk + 1) ? rc1 : rcown);
  double u0[TILE], u1[TILE];
#pragma unroll
  for (int tj = B; tj < TILE; ++tj) {
    u0[tj] = r0[tj * G + tc];
    u1[tj] = fma(-l10, u0[tj], r1[tj * G + tc]);
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
    for (int tj = B; tj < TILE; ++tj) a[ti][tj] = fma(-l1, u1[tj], fma(-l0, u0[tj], a[ti][tj]));
  }
  if (kk + 2 < G) {
    publish_pair<TILE, B>(a, kk + 2, buf ^ 1, lbuf, ubuf, rcs, tr, tc);
  } else if constexpr (B + 1 < TILE) {
    publish_pair<TILE, B + 1>(a, 0, buf ^ 1, lbuf, ubuf, rcs, tr, tc);
  }
}

template <int TILE, int B>
__device__ inline void dual_blocks(double (&a)[TILE][TILE], double (&b)[TILE][TILE], double* sa, double* sb, int tr,
                                   int tc) {
  if constexpr (B < TILE) {
    constexpr int MP = G * TILE;
    double rca = 0.0, rcb = 0.0;
#pragma nounroll
    for (int kk = 0; kk < G; kk += 2) {
      pair_round<TILE, B>(a, kk, sa, sa + 4 * MP, sa + 8 * MP, tr, tc, rca);
      pair_round<TILE, B>(b, kk, sb, sb + 4 * MP, sb + 8 * MP, tr, tc, rcb);
      __syncthreads();
    }
    scale_column<TILE, B>(a, rca, tr, tc);
    scale_column<TILE, B>(b, rcb, tr, tc);
    dual_blocks<TILE, B + 1>(a, b, sa, sb, tr, tc);
  }
}

// Two pair-scheme buffer sets of 9*MP doubles each.
template <int TILE>
__device__ inline void factor_dual(double (&a)[TILE][TILE], double (&b)[TILE][TILE], double* smem) {
  const int tid = threadIdx.x, tr = tid / G, tc = tid % G;
  constexpr int MP = G * TILE;
  double* sa = smem;
  double* sb = smem + 9 * MP;
  publish_pair<TILE, 0>(a, 0, 0, sa, sa + 4 * MP, sa + 8 * MP, tr, tc);
  publish_pair<TILE, 0>(b, 0, 0, sb, sb + 4 * MP, sb + 8 * MP, tr, tc);
  __syncthreads();
  dual_blocks<TILE, 0>(a, b, sa, sb, tr, tc);
}

// ------------------------------------------------------------- harness
template <int TILE>
__device__ inline void fill(double (&a)[TILE][TILE], int tr, int tc, unsigned seed) {
#pragma unroll
  for (int ti = 0; ti < TILE; ++ti)
#pragma unroll
    for (int tj = 0; tj < TILE; ++tj) {
      const unsigned i = ti * G + tr, j = tj * G + tc;
      unsigned h = (i * 2654435761u) ^ (j * 40503u) ^ seed;
      h ^= h >> 13;
      h *= 0x5bd1e995u;
      h ^= h >> 15;
      const double r = ((double)(h & 0xFFFF) / 65536.0) - 0.5;
      a[ti][tj] = (i == j) ? (2.0 * G * TILE + r) : r;
    }
}

template <int TILE, int V>
__global__ __launch_bounds__(NT) void k_factor_probe(int reps, double* out, unsigned long long* cyc,
                                                     unsigned seed_off) {
  extern __shared__ double smem[];
  PlanDev P{};
  P.tile = TILE;
  P.mpad = G * TILE;
  const EngineLds L = carve(smem, P);
  const int tid = threadIdx.x, tr = tid / G, tc = tid % G;
  double a[TILE][TILE];
  double acc = 0.0;
  unsigned long long total = 0;
  for (int r = 0; r < reps; ++r) {
    fill<TILE>(a, tr, tc, blockIdx.x * 7919u + r + seed_off);
    __syncthreads();
    const unsigned long long t0 = __builtin_amdgcn_s_memtime();
    if constexpr (V == 0)
      core_factor<TILE>(a, L);
    else
      factor_variant<TILE, V>(a, smem);
    __syncthreads();
    total += __builtin_amdgcn_s_memtime() - t0;
#pragma unroll
    for (int ti = 0; ti < TILE; ++ti)
#pragma unroll
      for (int tj = 0; tj < TILE; ++tj) acc += a[ti][tj];
  }
#pragma unroll
  for (int ti = 0; ti < TILE; ++ti)
#pragma unroll
    for (int tj = 0; tj < TILE; ++tj) out[((size_t)blockIdx.x * TILE * TILE + ti * TILE + tj) * NT + tid] = a[ti][tj];
  if (tid == 0) cyc[blockIdx.x] = total;
  if (acc == 12345.678) out[0] = acc;   // keeps the factorizations live
}


// Two cores per workgroup and repetition: a (seed as variant 0's) and b
// (seed + DUAL_OFF); both final tiles are written, a then b.
constexpr unsigned DUAL_OFF = 1u << 20;

template <int TILE>
__global__ __launch_bounds__(NT) void k_factor_dual(int reps, double* out, unsigned long long* cyc) {
  extern __shared__ double smem[];
  const int tid = threadIdx.x, tr = tid / G, tc = tid % G;
  double a[TILE][TILE], b[TILE][TILE];
  double acc = 0.0;
  unsigned long long total = 0;
  for (int r = 0; r < reps; ++r) {
    fill<TILE>(a, tr, tc, blockIdx.x * 7919u + r);
    fill<TILE>(b, tr, tc, blockIdx.x * 7919u + r + DUAL_OFF);
    __syncthreads();
    const unsigned long long t0 = __builtin_amdgcn_s_memtime();
    factor_dual<TILE>(a, b, smem);
    __syncthreads();
    total += __builtin_amdgcn_s_memtime() - t0;
#pragma unroll
    for (int ti = 0; ti < TILE; ++ti)
#pragma unroll
      for (int tj = 0; tj < TILE; ++tj) acc += a[ti][tj] + b[ti][tj];
  }
  const size_t half = (size_t)gridDim.x * TILE * TILE * NT;
#pragma unroll
  for (int ti = 0; ti < TILE; ++ti)
#pragma unroll
    for (int tj = 0; tj < TILE; ++tj) {
      const size_t o = ((size_t)blockIdx.x * TILE * TILE + ti * TILE + tj) * NT + tid;
      out[o] = a[ti][tj];
      out[half + o] = b[ti][tj];
    }
  if (tid == 0) cyc[blockIdx.x] = total;
  if (acc == 12345.678) out[0] = acc;
}

// Latency floors of one pivot-pair round, 72 rounds per repetition.
template <int S>
__global__ __launch_bounds__(NT) void k_floor(int reps, double* out, unsigned long long* cyc) {
  extern __shared__ double smem[];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  double x = 1.0 + tid * 1e-3;
  if (tid < 8) smem[tid] = 1.0;
  __syncthreads();
  unsigned long long total = 0;
  for (int r = 0; r < reps; ++r) {
    const unsigned long long t0 = __builtin_amdgcn_s_memtime();
#pragma nounroll
    for (int k = 0; k < 72; ++k) {
      if constexpr (S == 1) {
        __syncthreads();
      } else if constexpr (S == 2 || S == 3) {
        const int b = (k & 1) * 4;
        if (lane == 0) smem[b + wave] = x;
        __syncthreads();
        double v = (smem[b] + smem[b + 1]) + (smem[b + 2] + smem[b + 3]);
        if constexpr (S == 3) v = pivot_rcp(v);
        x = x * 0.5 + v * 1e-3;
      } else if constexpr (S == 4) {
#pragma unroll
        for (int i = 0; i < 8; ++i) x = fma(x, 0.999, 1e-3);
      } else {
        x = pivot_rcp(x) + 0.5;
      }
    }
    __syncthreads();
    total += __builtin_amdgcn_s_memtime() - t0;
  }
  out[blockIdx.x * NT + tid] = x;
  if (tid == 0) cyc[blockIdx.x] = total;
}

#define CK(x)                                                                          \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);            \
      exit(1);                                                                         \
    }                                                                                  \
  } while (0)

struct Result {
  double us_per_factor;   // kernel time / factorizations per CU
  double ticks;           // clock ticks per factorization (each workgroup's own)
};

template <int TILE, int V>
Result run_factor(int cus, int per_cu, int reps, size_t lds, std::vector<double>& host, unsigned seed_off = 0) {
  const int blocks = cus * per_cu;
  const size_t nout = (size_t)blocks * TILE * TILE * NT;
  double* out;
  unsigned long long* cyc;
  CK(hipMalloc(&out, nout * sizeof(double)));
  CK(hipMalloc(&cyc, blocks * sizeof(unsigned long long)));
  CK(hipFuncSetAttribute((const void*)k_factor_probe<TILE, V>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL((k_factor_probe<TILE, V>), dim3(blocks), dim3(NT), lds, 0, 1, out, cyc, seed_off);
  CK(hipGetLastError());
  CK(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  CK(hipEventRecord(e0));
  hipLaunchKernelGGL((k_factor_probe<TILE, V>), dim3(blocks), dim3(NT), lds, 0, reps, out, cyc, seed_off);
  CK(hipGetLastError());
  CK(hipEventRecord(e1));
  CK(hipDeviceSynchronize());
  float ms = 0;
  CK(hipEventElapsedTime(&ms, e0, e1));
  std::vector<unsigned long long> c(blocks);
  CK(hipMemcpy(c.data(), cyc, blocks * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  host.resize(nout);
  CK(hipMemcpy(host.data(), out, nout * sizeof(double), hipMemcpyDeviceToHost));
  CK(hipFree(out));
  CK(hipFree(cyc));
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
  double s = 0;
  for (auto x : c) s += (double)x;
  return {1e3 * ms / reps / per_cu, s / blocks / reps};
}

template <int S>
double run_floor(int cus, int reps) {
  const size_t lds = 130 * 1024;
  double* out;
  unsigned long long* cyc;
  CK(hipMalloc(&out, (size_t)cus * NT * sizeof(double)));
  CK(hipMalloc(&cyc, cus * sizeof(unsigned long long)));
  CK(hipFuncSetAttribute((const void*)k_floor<S>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL((k_floor<S>), dim3(cus), dim3(NT), lds, 0, reps, out, cyc);
  CK(hipGetLastError());
  CK(hipDeviceSynchronize());
  std::vector<unsigned long long> c(cus);
  CK(hipMemcpy(c.data(), cyc, cus * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  CK(hipFree(out));
  CK(hipFree(cyc));
  double s = 0;
  for (auto x : c) s += (double)x;
  return s / cus / reps / 72.0;
}

size_t differ(const std::vector<double>& x, const std::vector<double>& y) {
  size_t d = 0;
  for (size_t i = 0; i < x.size(); ++i) d += memcmp(&x[i], &y[i], sizeof(double)) != 0;
  return d;
}

template <int TILE, int V>
void variant_line(int cus, int reps, const std::vector<double>& ref, const char* name) {
  std::vector<double> h;
  const Result r = run_factor<TILE, V>(cus, 1, reps, 130 * 1024, h);
  printf("| %d | %d %s | %.2f | %.0f | %zu of %zu |\n", TILE, V, name, r.us_per_factor, r.ticks, differ(ref, h),
         ref.size());
  fflush(stdout);
}


// Two cores per workgroup (variant 6): time per factorization, and both
// cores checked against variant 0 on the same matrices.
template <int TILE>
void dual_line(int cus, int reps, const std::vector<double>& ref) {
  const size_t lds = 130 * 1024;
  const size_t half = (size_t)cus * TILE * TILE * NT;
  double* out;
  unsigned long long* cyc;
  CK(hipMalloc(&out, 2 * half * sizeof(double)));
  CK(hipMalloc(&cyc, cus * sizeof(unsigned long long)));
  CK(hipFuncSetAttribute((const void*)k_factor_dual<TILE>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL((k_factor_dual<TILE>), dim3(cus), dim3(NT), lds, 0, 1, out, cyc);
  CK(hipGetLastError());
  CK(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  CK(hipEventRecord(e0));
  hipLaunchKernelGGL((k_factor_dual<TILE>), dim3(cus), dim3(NT), lds, 0, reps, out, cyc);
  CK(hipGetLastError());
  CK(hipEventRecord(e1));
  CK(hipDeviceSynchronize());
  float ms = 0;
  CK(hipEventElapsedTime(&ms, e0, e1));
  std::vector<double> h(2 * half);
  CK(hipMemcpy(h.data(), out, 2 * half * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<unsigned long long> c(cus);
  CK(hipMemcpy(c.data(), cyc, cus * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  CK(hipFree(out));
  CK(hipFree(cyc));
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
  std::vector<double> refb;
  run_factor<TILE, 0>(cus, 1, reps, lds, refb, DUAL_OFF);
  const std::vector<double> ha(h.begin(), h.begin() + half), hb(h.begin() + half, h.end());
  double t = 0;
  for (auto x : c) t += (double)x;
  printf("| %d | 6 dual (two cores, shared barriers) | %.2f | %.0f | %zu + %zu of %zu |\n", TILE, 1e3 * ms / reps / 2,
         t / cus / reps / 2, differ(ref, ha), differ(refb, hb), 2 * half);
  fflush(stdout);
}

template <int TILE>
void tile_table(int cus, int reps) {
  std::vector<double> ref;
  const Result r0 = run_factor<TILE, 0>(cus, 1, reps, 130 * 1024, ref);
  printf("| %d | 0 library | %.2f | %.0f | reference |\n", TILE, r0.us_per_factor, r0.ticks);
  fflush(stdout);
  variant_line<TILE, 1>(cus, reps, ref, "deferred");
  variant_line<TILE, 2>(cus, reps, ref, "lookahead");
  variant_line<TILE, 3>(cus, reps, ref, "publish-first");
  variant_line<TILE, 4>(cus, reps, ref, "blocked");
  variant_line<TILE, 5>(cus, reps, ref, "quad");
  dual_line<TILE>(cus, reps, ref);
}

template <int TILE>
void occupancy_line(int cus, int per_cu, int reps) {
  std::vector<double> h;
  const size_t lds = (160 * 1024) / per_cu - 2048;   // admits per_cu workgroups per CU
  const Result r = run_factor<TILE, 0>(cus, per_cu, reps, lds, h);
  printf("| %d | %d | %.2f | %.0f |\n", TILE, per_cu, r.us_per_factor, r.ticks);
  fflush(stdout);
}

}  // namespace probe

int main(int argc, char** argv) {
  using namespace probe;
  const int reps = argc > 1 ? atoi(argv[1]) : 200;
  int dev = 0, cus = 0;
  CK(hipGetDevice(&dev));
  CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  printf("core factor probe: %d CUs, %d factorizations per workgroup\n\n", cus, reps);
  printf("| TILE | variant | us per factorization per CU | clock ticks | entries differing from the library's |\n");
  printf("|---:|---|---:|---:|---|\n");
  tile_table<9>(cus, reps);
  tile_table<5>(cus, reps);
  tile_table<2>(cus, reps);
  printf("\n| TILE | workgroups per CU | us per factorization per CU | clock ticks per factorization |\n");
  printf("|---:|---:|---:|---:|\n");
  for (int k = 1; k <= 3; ++k) occupancy_line<5>(cus, k, reps);
  for (int k = 1; k <= 4; ++k) occupancy_line<2>(cus, k, reps);
  printf("\n| floor of one round (72 rounds) | clock ticks per round |\n|---|---:|\n");
  printf("| barrier | %.1f |\n", run_floor<1>(cus, 100));
  printf("| LDS write, barrier, LDS read | %.1f |\n", run_floor<2>(cus, 100));
  printf("| same + pivot reciprocal | %.1f |\n", run_floor<3>(cus, 100));
  printf("| 8 dependent fp64 fma | %.1f |\n", run_floor<4>(cus, 100));
  printf("| pivot reciprocal chain | %.1f |\n", run_floor<5>(cus, 100));
  unsigned long long to = 0;
  CK(hipMemcpyFromSymbol(&to, HIP_SYMBOL(g_spin_timeouts), sizeof(to)));
  printf("\nbounded waits that timed out (variant 2): %llu\n", to);
  return 0;
}
