// The batched persistent kernel held to two waves per SIMD (at most 256
// registers per lane), for register tiles up to OCC2_MAX_TILE, with the
// plan's tables in memory: two workgroups, two cells, share a CU when their
// LDS fits, and each hides the other's latency.
#include "batched_kernel.hpp"

namespace fode {

int launch_batched_occ2(const BatchArgs& A, int blocks_per_cu, hipStream_t s) {
  const bool dae = A.p.constrained != 0 && A.p.loop != FODE_LOOP_BACKWARD_EULER;
  switch (A.P.tile) {
#define X(T)                                                               \
  case T:                                                                  \
    return dae ? launch_batched_as<T, true, false, 2>(A, blocks_per_cu, s) \
               : launch_batched_as<T, false, false, 2>(A, blocks_per_cu, s);
    X(1) X(2) X(3) X(4) X(5)
#undef X
  }
  return bfail("fode_batched_interval: no two-waves-per-SIMD kernel for this tile");
}

}  // namespace fode
