// Persistent split-step kernel: one launch runs K full steps of a plan.
//
// The pass kernels (col_pass_kernel, row_pass_kernel) cost one dispatch per
// sweep: at 512^2 a single instance's pass is one workgroup's critical path
// (8-10 us) plus the dispatch's ramp-up and drain. Here one grid of
// workgroups, all resident at once, walks a work list for the whole block
// of steps: for every sub-step, the column items (slot, component, column
// tile) and then the row items (slot, row block) of the plan's active slots,
// dealt round-robin to the workgroups in that order. Dependencies are per
// slot: a row item waits until every column item of its slot and sub-step
// is done, a column item until every row item of its slot's previous
// sub-step is done, each counted by a per-slot completion counter. The
// items themselves are the pass kernels' bodies (col_item, row_item), so the
// arithmetic, and therefore the result, is that of the pass kernels.
//
// Progress: every workgroup takes its items in the global order (sub-step,
// pass, item), and an item only waits on items earlier in that order, so
// with all workgroups resident the earliest unfinished item can always run
// (the host checks residency from the occupancy of the kernel before it
// enables the path). Every wait is bounded: past `wait_ticks` of the
// wall clock (100 MHz) the waiting workgroup sets a sticky abort flag, and
// every workgroup returns at its next wait once the flag is set, so the grid
// always drains; the host reports the abort as an error.
#pragma once

#include <hip/hip_ext.h>

#include "sgpe_kernels.h"

namespace sgpe {

// Completion counters of one slot within a launch (zeroed before it).
struct PersistSync {
  unsigned col_done;  // column items finished
  unsigned row_done;  // row items finished
};

// status[0]: sticky abort flag (a wait ran out of time); status[1]:
// workgroups finished in the current launch (zeroed before it).
enum { PST_ABORT = 0, PST_DONE = 1, PST_WORDS = 2 };

constexpr int kPersistMaxSlots = 256;

struct PersistArgs {
  const InstDev* insts;
  RunDev* runs;
  void* direct_psi;  // single-slot plans: the state's address (else null: per slot)
  int nslots;
  int n_steps;
  int d_first_prev;  // duration whose closing kinetic half-step is pending at launch
  int col_stream;
  int row_stream;
  const void* tw_x;  // row twiddles
  const void* tw_y;  // column twiddles
  PersistSync* sync;
  unsigned* status;
  long long wait_ticks;
};

// Every cross-workgroup datum of a launch (states, partial sums) is written
// and read with device-coherent accesses (CoherentBuf, st/ld_coherent), so
// the release of an item is: every wave waits for its writes to complete,
// then thread 0 counts the item as done; and a wait needs no acquire fence
// (the reads that follow are coherent themselves). Cache-wide writebacks
// and invalidations per item, what device-scope fences cost here, are
// avoided: with them (one per wave) an item took about 60 us.
__device__ __forceinline__ void persist_signal(unsigned* cnt) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Thread 0 waits until *cnt >= target (bounded, see above). Returns false if
// the launch aborts.
__device__ __forceinline__ bool persist_wait(const unsigned* cnt, unsigned target, unsigned* status,
                                             long long ticks, int* flag) {
  if (threadIdx.x == 0) {
    int ok = 1;
    const long long t0 = wall_clock64();
    while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(status + PST_ABORT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        ok = 0;
        break;
      }
      if (wall_clock64() - t0 > ticks) {
        __hip_atomic_store(status + PST_ABORT, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    *flag = ok;
  }
  __syncthreads();
  return *flag != 0;
}

// Zeroes the completion counters and the finished-workgroup count of a plan
// (a kernel, so that it is stream-ordered and graph-safe). The abort flag
// is kept: it is cleared only when the host rebinds the states.
static __global__ void persist_reset_kernel(PersistSync* __restrict__ sync, int n, unsigned* __restrict__ status,
                                            int clear_abort) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    sync[i].col_done = 0u;
    sync[i].row_done = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    status[PST_DONE] = 0u;
    if (clear_abort) status[PST_ABORT] = 0u;
  }
}

// Square grids N x N with separable operators and no or uniform coupling
// (the fast variants of both passes); both passes use blocks of the same
// size: RROWS rows of N / RE threads, CW columns of N / CE threads.
template <typename R, int N, int RE, int RROWS, int CE, int CW>
__global__ __launch_bounds__(RROWS*(N / RE)) void persist_steps_kernel(PersistArgs a) {
  constexpr int CM = CM_LOAD | CM_STORE;
  static_assert(RROWS * (N / RE) == CW * (N / CE), "row and column blocks must have the same size");
  typedef typename Cx<R>::T C;
  // The pass bodies' exchange buffers assume the alignment they have in the
  // pass kernels (dynamic LDS at offset 0): 16-byte elements and 128-byte
  // bank windows of the swizzled layouts. The static arrays below come
  // first, so the dynamic buffer is aligned explicitly.
  extern __shared__ __attribute__((aligned(256))) double smem[];
  __shared__ int s_list[kPersistMaxSlots];
  __shared__ int s_step0[kPersistMaxSlots];
  __shared__ unsigned char s_kspace[kPersistMaxSlots];
  __shared__ int s_nact;
  __shared__ int s_flag;

  // The active slots, in slot order, with their form and step count at
  // launch (read once: the last workgroup to finish updates them).
  if (threadIdx.x < 64) {
    int base = 0;
    for (int i0 = 0; i0 < a.nslots; i0 += 64) {
      const int i = i0 + (int)threadIdx.x;
      int act = 0, form = 0, step = 0;
      if (i < a.nslots) {
        act = a.runs[i].active;
        form = a.runs[i].form;
        step = a.runs[i].step;
      }
      const unsigned long long m = __ballot(act);
      const int pos = base + __popcll(m & ((1ull << threadIdx.x) - 1ull));
      if (act) {
        s_list[pos] = i;
        s_kspace[pos] = form == FORM_KSPACE ? 1 : 0;
        s_step0[pos] = step;
      }
      base += __popcll(m);
    }
    if (threadIdx.x == 0) s_nact = base;
  }
  __syncthreads();
  const int nact = s_nact;
  constexpr int kTiles = N / CW;
  constexpr int kColPer = 2 * kTiles;
  constexpr int kRowPer = N / RROWS;
  const int n_col = nact * kColPer;
  const int n_row = nact * kRowPer;
  const C* twx = reinterpret_cast<const C*>(a.tw_x);
  const C* twy = reinterpret_cast<const C*>(a.tw_y);

  for (int k = 0; k < a.n_steps; ++k) {
    for (int j = 0; j < 3; ++j) {
      const int sub = 3 * k + j;
      const int dur = j == 1 ? DUR_IN : DUR_OUT;
      const int dprev = sub == 0 ? a.d_first_prev : (j == 2 ? DUR_IN : DUR_OUT);
      for (int it = blockIdx.x; it < n_col; it += gridDim.x) {
        const int li = it / kColPer;
        const int r = it - li * kColPer;
        // uniform: the slot's descriptor is read with scalar loads
        const int inst = __builtin_amdgcn_readfirstlane(s_list[li]);
        if (sub > 0 && !persist_wait(&a.sync[inst].row_done, (unsigned)(sub * kRowPer), a.status, a.wait_ticks,
                                     &s_flag))
          return;
        const InstDev& P = a.insts[inst];
        C* psi = reinterpret_cast<C*>(a.direct_psi ? a.direct_psi : P.psi);
        const int kspace = sub == 0 ? s_kspace[li] : 0;
        const int comp = r / kTiles;
        col_item<R, N, CE, CW, false, CM>(P, psi, N, twy, COL_MID, dprev, dur, a.col_stream, comp,
                                      xcd_tile(r - comp * kTiles, kTiles), smem, [&]() -> int { return kspace; });
        persist_signal(&a.sync[inst].col_done);
      }
      for (int it = blockIdx.x; it < n_row; it += gridDim.x) {
        const int li = it / kRowPer;
        const int blk = it - li * kRowPer;
        const int inst = __builtin_amdgcn_readfirstlane(s_list[li]);
        if (!persist_wait(&a.sync[inst].col_done, (unsigned)((sub + 1) * kColPer), a.status, a.wait_ticks, &s_flag))
          return;
        const InstDev& P = a.insts[inst];
        C* psi = reinterpret_cast<C*>(a.direct_psi ? a.direct_psi : P.psi);
        const RunDev& run = a.runs[inst];
        row_item<R, N, RE, RROWS, false, CM>(P, psi, N, twx, dur, CW, a.row_stream, blk, smem, [&]() -> bool {
          // The first sub-step of a full step records the previous step's
          // populations (row block 0 of the slot).
          if (j == 0 && blk == 0) record_pops<(CM & CM_LOAD) != 0>(P, run.pops, run.pops_cap, s_step0[li] + k, kColPer);
          return true;
        });
        persist_signal(&a.sync[inst].row_done);
      }
    }
  }

  // The last workgroup to finish publishes the run state: the states are
  // hybrid and n_steps more full steps have started.
  __syncthreads();
  if (threadIdx.x == 0)
    s_flag = __hip_atomic_fetch_add(a.status + PST_DONE, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
             gridDim.x - 1;
  __syncthreads();
  if (s_flag && a.n_steps > 0) {
    for (int i = threadIdx.x; i < nact; i += blockDim.x) {
      a.runs[s_list[i]].form = FORM_HYBRID;
      a.runs[s_list[i]].step = s_step0[i] + a.n_steps;
    }
  }
}

typedef void (*PFn)(dim3, dim3, size_t, hipStream_t, hipEvent_t, hipEvent_t, const PersistArgs&);

template <typename R, int N, int RE, int RROWS, int CE, int CW>
void persist_launch(dim3 g, dim3 b, size_t l, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const PersistArgs& a) {
  if (e0)
    hipExtLaunchKernelGGL((persist_steps_kernel<R, N, RE, RROWS, CE, CW>), g, b, (uint32_t)l, s, e0, e1, 0, a);
  else
    hipLaunchKernelGGL((persist_steps_kernel<R, N, RE, RROWS, CE, CW>), g, b, l, s, a);
}

struct PEntry {
  int prec, n, row_e, row_rows, col_e, col_w;
  void* launch;
  const void* kernel;
};

struct PTable {
  const PEntry* entries;
  int count;
};

PTable persist_registry_f64();
PTable persist_registry_f32();

}  // namespace sgpe

#define SGPE_PERSIST_E(R, PR, N, RE, RR, CE, CW)                                                   \
  {PR, N, RE, RR, CE, CW, (void*)&sgpe::persist_launch<R, N, RE, RR, CE, CW>,                      \
   (const void*)&sgpe::persist_steps_kernel<R, N, RE, RR, CE, CW>},

// Compiled persistent shapes: (N, row E, rows per block, column E, columns
// per block) pairs whose row and column blocks have the same size.
#define SGPE_PERSIST_SHAPES(X, R, PR)                                                                 \
  X(R, PR, 128, 4, 4, 8, 8) X(R, PR, 128, 2, 2, 4, 4) X(R, PR, 256, 2, 1, 8, 4) X(R, PR, 256, 4, 2, 8, 4) \
  X(R, PR, 512, 4, 2, 8, 4) X(R, PR, 512, 2, 1, 8, 4) X(R, PR, 512, 4, 2, 4, 2) X(R, PR, 512, 2, 1, 4, 2) \
  X(R, PR, 512, 4, 1, 8, 2) X(R, PR, 512, 8, 2, 8, 2) X(R, PR, 1024, 4, 1, 8, 2) X(R, PR, 1024, 4, 1, 16, 4)

#define SGPE_DEFINE_PERSIST_REGISTRY(FN, R, PR)                               \
  namespace sgpe {                                                            \
  static const PEntry k_persist[] = {SGPE_PERSIST_SHAPES(SGPE_PERSIST_E, R, PR)}; \
  PTable FN() { return {k_persist, (int)(sizeof(k_persist) / sizeof(PEntry))}; } \
  }
