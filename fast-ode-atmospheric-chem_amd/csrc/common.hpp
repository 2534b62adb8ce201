// Device-side descriptions of a mechanism and of a symbolic LU plan, shared
// by every kernel of the library (gfx950, wave64, FP64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fastode_hip.h"

namespace fode {

constexpr int G = 16;
constexpr int NT = G * G;   // threads of one cell engine
constexpr double EPS = 2.220446049250313e-16;   // float64 machine epsilon
constexpr uint32_t NO_DIV = 0xFFFFu;

// Level-scheduled task tables and the Schur lists of a plan.  The pointers
// address global memory (plain kernels) or a workgroup's LDS copy (the
// persistent batched kernel copies them once per workgroup).
struct Tables {
  const uint2* f_tasks; const int* f_ptr; const uint32_t* f_pairs;
  const uint2* sf_tasks; const int* sf_ptr;
  const uint2* sb_tasks; const int* sb_ptr;
  const uint32_t* s_pairs;
  const uint32_t* sch_pairs;   // Schur pairs of the core entries, per thread
};

struct PlanDev {
  int n, c, m, tile, ns, nsch, nnz;
  int lds_len, lu_len, mpad, slot_stride, vec;   // vec = ns
  int f_nst, sf_nst, sb_nst;
  // sizes (tasks, pairs) of the tables, for staging them into LDS
  int n_f_tasks, n_f_pairs, n_sf_tasks, n_sb_tasks, n_s_pairs, n_sch_pairs;
  Tables t;                    // global-memory tables
  const int* sp_src; const int* sp_diag; const int* sp_row;
  const short* core_src;       // [NT][slot_stride] Jacobian entry, -1 fill, -2 identity
  const uint32_t* sch_mask;    // [NT][4] bit s: slot s has a Schur sum
  const int* sch_base;         // [NT + 1] Schur pair range of each thread
  int sch_depth;               // most Schur sums of one thread
  const int* core_row;         // [mpad] old row of core row, -1 padding
  const int* perm;             // new -> old
};

struct MechDev {
  int n, R, P, nnz, ncon;
  const int* react; const int* rhs_ptr; const int* rhs_rxn; const double* rhs_coef;
  const int* jac_ptr; const int* jac_src; const double* jac_coef; const double* jac_const;
  const int* con_rows; const double* con_coef; const double* con_total;
  const int* con_of_row;   // n: constraint index of a row or -1
  // transposes built by fode_mech_create (for the adjoint):
  const int* st_ptr; const int* st_sp; const double* st_coef;   // S^T by reaction (R+1)
  const int* sp_ptr; const int* sp_jp;   // reactant slots j*P+p of each species (n+1)
};

}  // namespace fode
