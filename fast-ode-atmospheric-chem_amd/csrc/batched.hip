// Batched persistent form: one launch integrates every group of a batch over
// one host interval.
//
// Workgroups (256 threads) are persistent and take groups from an atomic
// queue.  A workgroup first copies the plan's task tables and Schur lists
// into its LDS (they are read by every attempt of every cell), then runs
// each group's whole step loop (the controller of Algorithm.solve,
// fastode/solvers.py:305-415, on the device) and performs every attempt cell
// by cell: the cell's rates, Jacobian, stage matrix and its LU, the stage
// solves and the stage combination all stay in LDS and registers; only the
// cell's state, rate constants and new state touch memory.  The group's error
// is the sum of the cells' partial sums in cell order, exactly as the plain
// form forms it, so both forms take the same steps.
#include "batched_kernel.hpp"

#include <mutex>

using namespace fode;

// Scratch of fode_batched_interval_noalloc: the work queue's counter in its
// first 256 bytes, then one slot of group_size * n doubles (two for a
// projected, constrained tableau) per persistent workgroup.  Each GPU the
// library runs on holds its own copy until the process ends.
#define FODE_NOALLOC_SCRATCH_BYTES (32u << 20)
static __device__ __align__(256) double scratch[FODE_NOALLOC_SCRATCH_BYTES / sizeof(double)];

#ifdef FODE_PHASE_TIMERS
// Cycle sums per phase over every workgroup of every timed launch.
static __device__ __align__(256) unsigned long long phase_cycles[PH_COUNT];
#endif

namespace {

__global__ void k_reset_queue(int* q) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *q = 0;
}

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of this GPU's scratch, looked up once per device with
// that device current.  The array's name in host code is only a stand-in for
// the symbol; kernels get this address.
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return bfail(std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= MAX_DEVICES) return bfail("fode_batched_interval_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return bfail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// Checks shared by the interval entry points; fills A but for work/queue.
int prepare(void* plan, void* mech, const fode_ctrl_params* p, const fode_batch_desc* b, const char* who,
            BatchArgs& A) {
  const std::string w(who);
  if (!plan || !mech || !p || !b) return bfail(w + ": null argument");
  if (b->C <= 0 || b->group_size <= 0 || b->ngroups != (b->C + b->group_size - 1) / b->group_size)
    return bfail(w + ": C, group_size and ngroups disagree");
  if (p->stages < 1 || p->stages > 4) return bfail(w + ": bad tableau");
  if (!b->Y || !b->k || !b->kmul || !b->atol || !b->mass || !b->g) return bfail(w + ": null array");
  if ((b->hlog != nullptr) != (b->nacc != nullptr) || (b->hlog && b->hcap < p->max_steps))
    return bfail(w + ": hlog and nacc go together, with hcap >= max_steps");
  if (p->loop < FODE_LOOP_ALGORITHM || p->loop > FODE_LOOP_BACKWARD_EULER) return bfail(w + ": unknown step loop");
  if (p->loop == FODE_LOOP_BACKWARD_EULER && (p->n_reductions < 0 || p->n_reductions > 8 || p->max_iterations < 1))
    return bfail(w + ": bad backward Euler parameters");
  if (p->loop != FODE_LOOP_ALGORITHM && b->hlog) return bfail(w + ": step records are kept for Algorithm's loop only");
  A.P = *static_cast<const PlanDev*>(fode_internal_plan(plan));
  A.M = *static_cast<const MechDev*>(fode_internal_mech(mech));
  if (A.P.n != A.M.n || A.P.nnz != A.M.nnz) return bfail(w + ": plan and mechanism differ");
  if (A.P.tile < 1 || A.P.tile > 10) return bfail(w + ": unsupported tile");
  A.p = *p;
  A.b = *b;
  A.lay = choose_layout(A.P, A.M, p->stages, b->options);
  A.work_by_block = 0;
  A.work_half = 0;
  A.max_blocks = 0;
  A.timers = nullptr;
#ifdef FODE_PHASE_TIMERS
  void* t = nullptr;
  if (hipGetSymbolAddress(&t, HIP_SYMBOL(phase_cycles)) != hipSuccess) return bfail(w + ": phase timers");
  A.timers = static_cast<unsigned long long*>(t);
#endif
  return 0;
}

// Two or more workgroups per CU asked for, on a small register tile: the
// instantiation held to 256 registers per lane, with the tables in memory
// (so that two cells' LDS fits one CU); otherwise one wave per SIMD.
int launch(const BatchArgs& A, int blocks_per_cu, hipStream_t s) {
  hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, s, A.b.queue);
  if (blocks_per_cu >= 2 && A.P.tile <= OCC2_MAX_TILE) {
    BatchArgs B = A;
    B.lay = make_layout(A.P, A.M, A.p.stages, STAGE_NONE);
    return launch_batched_occ2(B, blocks_per_cu, s);
  }
  return A.lay.tables > 0 ? launch_batched_tiles<true>(A, blocks_per_cu, s)
                          : launch_batched_unstaged(A, blocks_per_cu, s);
}

}  // namespace

extern "C" {

int fode_batched_lds_bytes(void* plan, void* mech, int stages, int64_t* bytes) {
  if (!plan || !mech || !bytes || stages < 1 || stages > 4) return bfail("fode_batched_lds_bytes: bad argument");
  const PlanDev& P = *static_cast<const PlanDev*>(fode_internal_plan(plan));
  const MechDev& M = *static_cast<const MechDev*>(fode_internal_mech(mech));
  *bytes = (int64_t)choose_layout(P, M, stages).total * (int64_t)sizeof(double);
  return 0;
}

struct GraphHandle {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int fode_batched_graph_create(void* plan, void* mech, const fode_ctrl_params* p,
                              const fode_batch_desc* b, int blocks_per_cu, void** out) {
  if (!out) return bfail("fode_batched_graph_create: null argument");
  hipStream_t s = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return bfail(std::string("hipStreamCreate: ") + hipGetErrorString(e));
  auto* h = new GraphHandle();
  int rc = 0;
  e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    rc = bfail(std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
  } else {
    rc = fode_batched_interval(plan, mech, p, b, blocks_per_cu, s);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (rc == 0 && e != hipSuccess) rc = bfail(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    h->graph = g;
    if (rc == 0) {
      e = hipGraphInstantiate(&h->exec, g, nullptr, nullptr, 0);
      if (e != hipSuccess) rc = bfail(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
  }
  (void)hipStreamDestroy(s);
  if (rc != 0) {
    fode_graph_destroy(h);
    return rc;
  }
  *out = h;
  return 0;
}

int fode_graph_launch(void* graph, void* stream) {
  auto* h = static_cast<GraphHandle*>(graph);
  if (!h || !h->exec) return bfail("fode_graph_launch: null graph");
  hipError_t e = hipGraphLaunch(h->exec, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return bfail(std::string("hipGraphLaunch: ") + hipGetErrorString(e));
  return 0;
}

int fode_graph_destroy(void* graph) {
  auto* h = static_cast<GraphHandle*>(graph);
  if (!h) return 0;
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  delete h;
  return 0;
}

int fode_batched_interval(void* plan, void* mech, const fode_ctrl_params* p, const fode_batch_desc* b,
                          int blocks_per_cu, void* stream) {
  BatchArgs A;
  if (prepare(plan, mech, p, b, "fode_batched_interval", A)) return -1;
  if (!b->work || !b->queue) return bfail("fode_batched_interval: null array");
  return launch(A, blocks_per_cu, reinterpret_cast<hipStream_t>(stream));
}

int fode_phase_timers(uint64_t* out, int n, int reset) {
#ifdef FODE_PHASE_TIMERS
  if (!out || n < 0) return bfail("fode_phase_timers: bad argument");
  unsigned long long host[PH_COUNT];
  hipError_t e = hipMemcpyFromSymbol(host, HIP_SYMBOL(phase_cycles), sizeof(host));
  if (e != hipSuccess) return bfail(std::string("fode_phase_timers: ") + hipGetErrorString(e));
  for (int i = 0; i < n && i < PH_COUNT; ++i) out[i] = host[i];
  if (reset) {
    const unsigned long long zero[PH_COUNT] = {};
    e = hipMemcpyToSymbol(HIP_SYMBOL(phase_cycles), zero, sizeof(zero));
    if (e != hipSuccess) return bfail(std::string("fode_phase_timers: ") + hipGetErrorString(e));
  }
  return PH_COUNT;
#else
  (void)out; (void)n; (void)reset;
  return bfail("fode_phase_timers: this library was built without FODE_PHASE_TIMERS");
#endif
}

int fode_batched_interval_noalloc(void* plan, void* mech, const fode_ctrl_params* p, const fode_batch_desc* b,
                                  int blocks_per_cu, void* stream) {
  BatchArgs A;
  if (prepare(plan, mech, p, b, "fode_batched_interval_noalloc", A)) return -1;
  if (b->work || b->queue)
    return bfail("fode_batched_interval_noalloc: work and queue must be NULL (the library's scratch holds them)");
  void* base = nullptr;
  if (scratch_address(&base)) return -1;
  const bool projected = p->constrained && p->loop != FODE_LOOP_BACKWARD_EULER;
  const long slot = (long)b->group_size * A.M.n;              // doubles per workgroup and half
  const long avail = (long)(FODE_NOALLOC_SCRATCH_BYTES / sizeof(double)) - 32;   // after the queue's 256 bytes
  const long slots = avail / (slot * (projected ? 2 : 1));
  if (slots < 1)
    return bfail("fode_batched_interval_noalloc: group_size * n = " + std::to_string(slot) + " doubles" +
                 (projected ? " (x2 with the projection)" : "") + " exceed the " +
                 std::to_string(FODE_NOALLOC_SCRATCH_BYTES) + "-byte scratch");
  A.b.queue = static_cast<int32_t*>(base);
  A.b.work = static_cast<double*>(base) + 32;
  A.work_by_block = 1;
  A.work_half = slots * slot;
  A.max_blocks = (int)std::min<long>(slots, 1L << 30);
  return launch(A, blocks_per_cu, reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
