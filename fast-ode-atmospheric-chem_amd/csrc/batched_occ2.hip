// The batched persistent kernel held to two waves per SIMD (at most 256
// registers per lane), for register tiles up to OCC2_MAX_TILE, with the
// plan's tables where they fit: two workgroups, two cells, share a CU and
// each hides the other's latency.  All tables go to LDS when the CU still
// holds blocks_per_cu workgroups with them, else the solves' tables only (a
// solve's level-scheduled stages then read their task lists from LDS instead
// of memory), else none; the arithmetic, and so the bits, are the same.
#include "batched_kernel.hpp"

namespace fode {

namespace {

// Workgroups of kernel kern that fit one CU with bytes of dynamic LDS.  A
// layout beyond the device's limit is 0 without asking the runtime, and an
// error of the query is cleared here, so that the launch's own
// hipGetLastError() does not report it.
int resident_per_cu(const void* kern, size_t bytes) {
  int dev = 0, maxlds = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (bytes + sizeof(fode_group_ctrl) > (size_t)maxlds) return 0;
  int occ = 0;
  if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, NT, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return occ;
}

// Whether blocks_per_cu workgroups of the STAGE kernel fit one CU; B.lay
// becomes that kernel's layout.
template <int TILE, bool DAE, int STAGE>
bool occ2_fits(BatchArgs& B, int blocks_per_cu) {
  const bool be = B.p.loop == FODE_LOOP_BACKWARD_EULER;
  B.lay = make_layout(B.P, B.M, B.p.stages, STAGE);
  const void* kern = be ? (const void*)k_batched_be<TILE, STAGE, 2> : (const void*)k_batched<TILE, DAE, STAGE, 2>;
  const int want = blocks_per_cu < 2 ? 2 : blocks_per_cu;
  return resident_per_cu(kern, (size_t)B.lay.total * sizeof(double)) >= want;
}

// Every table in LDS when that still fits blocks_per_cu workgroups on a CU,
// else the solves' tables, else none.
template <int TILE, bool DAE>
int launch_occ2_tile(const BatchArgs& A, int blocks_per_cu, hipStream_t s) {
  BatchArgs B = A;
  if (occ2_fits<TILE, DAE, STAGE_ALL>(B, blocks_per_cu))
    return launch_batched_as<TILE, DAE, STAGE_ALL, 2>(B, blocks_per_cu, s);
  if (occ2_fits<TILE, DAE, STAGE_SOLVE>(B, blocks_per_cu))
    return launch_batched_as<TILE, DAE, STAGE_SOLVE, 2>(B, blocks_per_cu, s);
  B.lay = make_layout(A.P, A.M, A.p.stages, STAGE_NONE);
  return launch_batched_as<TILE, DAE, STAGE_NONE, 2>(B, blocks_per_cu, s);
}

}  // namespace

int launch_batched_occ2(const BatchArgs& A, int blocks_per_cu, hipStream_t s) {
  const bool dae = A.p.constrained != 0 && A.p.loop != FODE_LOOP_BACKWARD_EULER;
  switch (A.P.tile) {
#define X(T)                                                         \
  case T:                                                            \
    return dae ? launch_occ2_tile<T, true>(A, blocks_per_cu, s)      \
               : launch_occ2_tile<T, false>(A, blocks_per_cu, s);
    X(1) X(2) X(3) X(4) X(5)
#undef X
  }
  return bfail("fode_batched_interval: no two-waves-per-SIMD kernel for this tile");
}

}  // namespace fode
