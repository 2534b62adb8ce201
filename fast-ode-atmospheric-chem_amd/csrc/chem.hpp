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
