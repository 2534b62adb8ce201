// The batched persistent kernel with the plan's tables read from memory
// (when they do not fit in LDS beside the cell, or the caller asks for it).
#include "batched_kernel.hpp"

namespace fode {

int launch_batched_unstaged(const BatchArgs& A, int blocks_per_cu, hipStream_t s) {
  return launch_batched_tiles<false>(A, blocks_per_cu, s);
}

}  // namespace fode
