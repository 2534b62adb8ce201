// Mass-action rates, forcing and Jacobian partials on the device
// (fastode/chem.py:81-102), shared by the plain and the batched kernels.
#pragma once
#include "common.hpp"

namespace fode {

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

// f + sum_q vals[src[q]] * coef[q] over q in [q, qe), added in q order (the
// same bits as one term at a time), with the table loads and gathers of TERMS
// terms in flight together: a long row (TS1 has species with 192 terms and a
// Jacobian entry with 110) costs one load latency per TERMS terms, not per
// term.
// WIDE (kernels with registers to spare): rounds of 2*TERMS terms first, so
// that a long row (TS1's longest forcing row has 192 terms) waits on half as
// many load latencies; the additions keep the term order, so the bits are
// those of the narrow rounds.
constexpr int TERMS = 8;
template <bool WIDE = false>
__device__ inline double ordered_dot(double f, const double* vals, const int* src, const double* coef, int q,
                                     int qe) {
  if constexpr (WIDE) {
    constexpr int W = 2 * TERMS;
    for (; q + W <= qe; q += W) {
      int x[W];
      double c[W], r[W];
#pragma unroll
      for (int u = 0; u < W; ++u) {
        x[u] = src[q + u];
        c[u] = coef[q + u];
      }
#pragma unroll
      for (int u = 0; u < W; ++u) r[u] = vals[x[u]];
#pragma unroll
      for (int u = 0; u < W; ++u) f = f + r[u] * c[u];
    }
  }
  for (; q + TERMS <= qe; q += TERMS) {
    int x[TERMS];
    double c[TERMS], r[TERMS];
#pragma unroll
    for (int u = 0; u < TERMS; ++u) {
      x[u] = src[q + u];
      c[u] = coef[q + u];
    }
#pragma unroll
    for (int u = 0; u < TERMS; ++u) r[u] = vals[x[u]];
#pragma unroll
    for (int u = 0; u < TERMS; ++u) f = f + r[u] * c[u];
  }
  if (q < qe) {
    // the last one to TERMS - 1 terms: loads in flight together (indices
    // clamped into the row, the extra terms discarded), added in order
    const int last = qe - 1;
    int x[TERMS - 1];
    double c[TERMS - 1], r[TERMS - 1];
#pragma unroll
    for (int u = 0; u < TERMS - 1; ++u) {
      const int qq = q + u < last ? q + u : last;
      x[u] = src[qq];
      c[u] = coef[qq];
    }
#pragma unroll
    for (int u = 0; u < TERMS - 1; ++u) r[u] = vals[x[u]];
#pragma unroll
    for (int u = 0; u < TERMS - 1; ++u)
      if (q + u < qe) f = f + r[u] * c[u];
  }
  return f;
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
  return ordered_dot(0.0, rates, M.rhs_rxn, M.rhs_coef, M.rhs_ptr[i], M.rhs_ptr[i + 1]);
}

}  // namespace fode
