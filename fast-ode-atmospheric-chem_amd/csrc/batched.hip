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
#include <hip/hip_runtime.h>

#include <string>

#include "cell.hpp"

using namespace fode;

// Shared with fastode_hip.hip, which owns the handles and the error text.
extern "C" const void* fode_internal_plan(void* plan);
extern "C" const void* fode_internal_mech(void* mech);
extern "C" void fode_internal_set_error(const char* msg);

namespace {

int bfail(const std::string& m) {
  fode_internal_set_error(m.c_str());
  return -1;
}

struct BatchArgs {
  PlanDev P;
  MechDev M;
  fode_ctrl_params p;
  fode_batch_desc b;
  Layout lay;
};

// Tables in LDS when everything fits in one workgroup's LDS, else in memory.
Layout choose_layout(const PlanDev& P, const MechDev& M, int S) {
  int dev = 0, maxlds = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  const Layout staged = make_layout(P, M, S, true);
  if ((size_t)staged.total * sizeof(double) + sizeof(fode_group_ctrl) <= (size_t)maxlds) return staged;
  return make_layout(P, M, S, false);
}

// One attempt on one cell. Returns the cell's error partial sum (thread 0).
template <int TILE>
__device__ double cell_attempt(const BatchArgs& A, const Tables& T, const EngineLds& L,
                               const CellLds& V, const fode_group_ctrl& g, int cell, double* ynew_out) {
  const MechDev& M = A.M;
  const int n = M.n, tid = threadIdx.x;
  const double* yg = A.b.Y + (size_t)cell * n;
  const double* kg = A.b.k + (size_t)cell * A.b.k_stride;
  for (int i = tid; i < n; i += NT) V.y[i] = yg[i];
  for (int j = tid; j < M.R; j += NT) V.kc[j] = kg[j] * A.b.kmul[j];
  __syncthreads();
  cell_stages<TILE>(A.P, T, M, A.p, A.b.mass, L, V, g.H, g.alpha, ynew_out);
  return cell_err_sum(V.y, V.w, V.fr, A.b.atol, A.b.rtol, n, V.red);
}

template <int TILE>
__global__ __launch_bounds__(NT) void k_batched(BatchArgs A) {
  extern __shared__ double smem[];
  const int n = A.M.n, S = A.p.stages;
  const Tables T = A.lay.tables > 0 ? stage_tables(A.P, reinterpret_cast<uint32_t*>(smem)) : A.P.t;
  const EngineLds L = carve(smem + A.lay.engine, A.P);
  const CellLds V = carve_cell(smem, A.lay, n, A.M.R, S);
  __shared__ fode_group_ctrl gs;
  const int tid = threadIdx.x;
  const fode_batch_desc& B = A.b;
  for (;;) {
    if (tid == 0) V.flags[0] = atomicAdd(B.queue, 1);
    __syncthreads();
    const int group = V.flags[0];
    __syncthreads();
    if (group >= B.ngroups) break;
    const int c0 = group * B.group_size;
    const int cells = min(B.group_size, B.C - c0);
    if (tid == 0) {
      gs = B.g[group];
      gs.dt = B.dt;
      ctrl_init(A.p, gs);
      ctrl_outer(A.p, gs);
    }
    __syncthreads();
    int nrec = 0;   // accepted steps recorded (thread 0)
    while (gs.action != FODE_FINISH) {
      double sum = 0.0;
      for (int c = 0; c < cells; ++c) {
        double* ynew = B.work + (size_t)(c0 + c) * n;
        const double cs = cell_attempt<TILE>(A, T, L, V, gs, c0 + c, ynew);
        sum = sum + cs;
      }
      if (tid == 0) {
        const double h_used = gs.H;
        ctrl_after(A.p, gs, err_from_sum(sum, (double)cells * n));
        if (gs.accepted && B.hlog && nrec < B.hcap) B.hlog[(size_t)group * B.hcap + nrec++] = h_used;
      }
      __syncthreads();
      if (gs.accepted) {
        const long N = (long)cells * n;
        double* Y = B.Y + (size_t)c0 * n;
        const double* W = B.work + (size_t)c0 * n;
        for (long i = tid; i < N; i += NT) Y[i] = W[i];
      }
      __syncthreads();
    }
    // clamp (fastode/solvers.py:89-97) and record the call's status
    {
      const long N = (long)cells * n;
      double* Y = B.Y + (size_t)c0 * n;
      for (long i = tid; i < N; i += NT) {
        if (A.p.constrained && B.mass[i % n] != 1.0) continue;
        Y[i] = fmax(Y[i], 0.0);
      }
    }
    if (tid == 0) {
      if (gs.status != FODE_CONVERGED) {
        if (gs.fails == 0) gs.first_fail = (int)gs.counts[FODE_C_CALLS] - 1;
        gs.fails += 1;
      }
      B.g[group] = gs;
      if (B.nacc) B.nacc[group] = nrec;
    }
    __syncthreads();
  }
}

__global__ void k_reset_queue(int* q) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *q = 0;
}

template <int TILE>
int launch_batched(BatchArgs A, int blocks_per_cu, hipStream_t s) {
  const size_t bytes = (size_t)A.lay.total * sizeof(double);
  int dev = 0, cus = 0, maxlds = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (bytes + sizeof(fode_group_ctrl) > (size_t)maxlds)
    return bfail("fode_batched_interval: the plan needs " + std::to_string(bytes) + " bytes of LDS, above " +
                 std::to_string(maxlds));
  hipError_t e = hipFuncSetAttribute((const void*)k_batched<TILE>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return bfail(std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  int occ = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_batched<TILE>, NT, bytes);
  if (e != hipSuccess || occ < 1) occ = 1;
  const int per_cu = blocks_per_cu > 0 ? blocks_per_cu : occ;
  int blocks = cus * per_cu;
  if (blocks > A.b.ngroups) blocks = A.b.ngroups;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, s, A.b.queue);
  hipLaunchKernelGGL(k_batched<TILE>, dim3(blocks), dim3(NT), bytes, s, A);
  e = hipGetLastError();
  if (e != hipSuccess) return bfail(std::string("k_batched: ") + hipGetErrorString(e));
  return 0;
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
  if (!plan || !mech || !p || !b) return bfail("fode_batched_interval: null argument");
  if (b->C <= 0 || b->group_size <= 0 || b->ngroups != (b->C + b->group_size - 1) / b->group_size)
    return bfail("fode_batched_interval: C, group_size and ngroups disagree");
  if (p->stages < 1 || p->stages > 4) return bfail("fode_batched_interval: bad tableau");
  if (!b->Y || !b->k || !b->kmul || !b->atol || !b->mass || !b->g || !b->work || !b->queue)
    return bfail("fode_batched_interval: null array");
  if ((b->hlog != nullptr) != (b->nacc != nullptr) || (b->hlog && b->hcap < p->max_steps))
    return bfail("fode_batched_interval: hlog and nacc go together, with hcap >= max_steps");
  BatchArgs A;
  A.P = *static_cast<const PlanDev*>(fode_internal_plan(plan));
  A.M = *static_cast<const MechDev*>(fode_internal_mech(mech));
  if (A.P.n != A.M.n || A.P.nnz != A.M.nnz) return bfail("fode_batched_interval: plan and mechanism differ");
  A.p = *p;
  A.b = *b;
  A.lay = choose_layout(A.P, A.M, p->stages);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  switch (A.P.tile) {
#define X(T) case T: return launch_batched<T>(A, blocks_per_cu, s);
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10)
#undef X
  }
  return bfail("fode_batched_interval: unsupported tile");
}

}  // extern "C"
