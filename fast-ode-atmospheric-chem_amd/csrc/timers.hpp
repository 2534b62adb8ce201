// Phase timers of the batched kernel, a build option (-DFODE_PHASE_TIMERS,
// built as libfastode_hip_timed.so): thread 0 of each workgroup reads the
// shader clock at the end of every phase of a cell-attempt and charges the
// cycles since the previous mark to that phase; at the end of the launch the
// workgroup's sums are added to a device array that fode_phase_timers reads.
// Without the option the marks compile to nothing.
#pragma once
#include <hip/hip_runtime.h>

namespace fode {

enum {
  PH_LOAD = 0,        // cell state and rate constants into LDS
  PH_RATES,           // rates and forcing F(Y)
  PH_JACOBIAN,        // rate partials and Jacobian values
  PH_ASSEMBLE,        // alpha M - J into the sparse entries and the core tile
  PH_SPARSE_FACTOR,   // level-scheduled sparse factor tasks
  PH_SCHUR,           // the core's update from the sparse pivots
  PH_CORE_FACTOR,     // register-tiled dense core LU
  PH_INVERSES,        // inverses of the core's diagonal 16x16 blocks
  PH_STAGE_SETUP,     // stage points, their forcing, right-hand sides, stage vectors
  PH_SPARSE_SOLVE,    // level-scheduled forward and back substitution
  PH_CORE_SOLVE,      // block substitution of the core
  PH_COMBINE,         // new state, error estimate, the cell's error sum
  PH_CONTROL,         // group control, acceptance copies, queue
  PH_COUNT
};

#ifdef FODE_PHASE_TIMERS
// [0, PH_COUNT): cycles per phase; [PH_COUNT]: the last mark
__shared__ unsigned long long fode_phase_acc[PH_COUNT + 1];

__device__ inline void phase_start() {
  if (threadIdx.x == 0) {
    for (int i = 0; i < PH_COUNT; ++i) fode_phase_acc[i] = 0;
    fode_phase_acc[PH_COUNT] = clock64();
  }
}
__device__ inline void phase_mark(int phase) {
  if (threadIdx.x == 0) {
    const unsigned long long t = clock64();
    fode_phase_acc[phase] += t - fode_phase_acc[PH_COUNT];
    fode_phase_acc[PH_COUNT] = t;
  }
}
__device__ inline void phase_flush(unsigned long long* out) {
  if (threadIdx.x == 0 && out)
    for (int i = 0; i < PH_COUNT; ++i) atomicAdd(out + i, fode_phase_acc[i]);
}
#define FODE_PHASE(p) ::fode::phase_mark(p)
#else
__device__ inline void phase_start() {}
__device__ inline void phase_flush(unsigned long long*) {}
#define FODE_PHASE(p) ((void)0)
#endif

}  // namespace fode
