// Task-list kernels of the PD iteration (see tasks.hpp).  Each block finds
// its task, checks the task's device-side predicates (uniform per block, so a
// whole block returns together) and runs the shared block / voxel bodies.
#include <type_traits>

#include "cart_fused.hpp"
#include "reg_device.hpp"
#include "tasks.hpp"

namespace avionic {

enum Family { kVox = 0, kCol = 1, kRow = 2, kDp1 = 3, kDp2 = 4 };

template <int F>
__device__ inline int first_block(const Task &t) {
  switch (F) {
    case kVox: return t.b_vox;
    case kCol: return t.b_col;
    case kRow: return t.b_row;
    case kDp1: return t.b_dp1;
    default: return t.b_dp2;
  }
}

// Last task whose first block does not exceed b (tasks are in block order;
// a task with no blocks in this family shares its start with the next one):
// a binary search over the non-decreasing first blocks.
template <int F>
__device__ inline int find_task(const Task *__restrict__ tasks, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first_block<F>(tasks[mid]) <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Renumbering blocks so that every XCD walks one contiguous range keeps the
// neighbouring frames that the step-rule reduction reads in one XCD's L2 (a
// bijection on [0, n)).  Measured on MI355X it helps that kernel and slows
// the streaming dual / primal updates, which keep the hardware order.
__device__ inline int xcd_block(int b, int n) {
  constexpr int X = 8;
  const int q = n / X, r = n % X;
  const int x = b % X, k = b / X;
  return x * q + min(x, r) + k;
}

__device__ inline bool task_active(const DevScalars *sc) { return sc->iter < sc->maxit; }
__device__ inline bool task_adapt(const DevScalars *sc) {
  const int k = sc->iter;
  return sc->schedule == 1 || (sc->schedule == 0 && (k < 10 || (k % 50) == 0));
}

// Dual and primal updates over the blocks of family F: kDp2 runs a pair of
// voxels of a row per lane (tasks of even width), kDp1 (tasks of odd width)
// and kVox (every task) one voxel per lane.
template <int F>
__device__ inline bool lane_voxels(const Task &T, int blk, Vox &v) {
  return vox_of_block<false>(T.g, blk - first_block<F>(T), threadIdx.x, v);
}
template <int F>
__device__ inline bool lane_voxels(const Task &T, int blk, Vox2 &v) {
  return vox2_of_block(T.g, blk - T.b_dp2, threadIdx.x, v);
}

template <int M, typename DT, int F>
__global__ __launch_bounds__(RT) void kt_dual(const Task *__restrict__ tasks, int n) {
  typedef typename std::conditional<F == kDp2, Vox2, Vox>::type V;
  const int blk = blockIdx.x;
  const Task &T = tasks[find_task<F>(tasks, n, blk)];
  if (!task_active(T.sc)) return;
  V v;
  if (!lane_voxels<F>(T, blk, v)) return;
  dual_voxel<M, DT>(T.g, T.rp, T.ext, (DT *)T.y, (float)T.sc->sigma, v);
}

template <int M, typename DT, int F>
__global__ __launch_bounds__(RT) void kt_primal(const Task *__restrict__ tasks, int n) {
  typedef typename std::conditional<F == kDp2, Vox2, Vox>::type V;
  const int blk = blockIdx.x;
  const Task &T = tasks[find_task<F>(tasks, n, blk)];
  if (!task_active(T.sc)) return;
  V v;
  if (!lane_voxels<F>(T, blk, v)) return;
  primal_voxel<M, DT>(T.g, T.rp, T.x, T.ext, (const DT *)T.y, T.khz, (float)T.sc->tau, v);
}

template <int M>
__global__ __launch_bounds__(RT) void kt_step_reg(const Task *__restrict__ tasks, int n) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const Task &T = tasks[find_task<kVox>(tasks, n, blk)];
  if (!task_active(T.sc) || !task_adapt(T.sc)) return;
  const int lb = blk - T.b_vox;
  Vox v;
  double ak = 0.0, ax = 0.0;
  if (vox_of_block<true>(T.g, lb, threadIdx.x, v)) step_voxel<M>(T.g, T.rp, T.ext, T.x, v, ak, ax);
  block_reduce2(ak, ax, T.part_reg + 2 * lb);
}

// Column passes, one kernel per kind of column plan (col_kind): kind 0 is the
// runtime plan, kind 1 the compile-time plans (a switch over the plan; the
// centring flag stays a runtime value: one body per plan and centring needed
// 2-3 times the VGPRs, and one kernel per plan served mixed sizes 6% slower
// than one kernel for all of them).  A launch covers its kind's span of
// blocks and is made only when such a task is resident.
template <bool DELTA, class P>
__device__ __forceinline__ void col_forward_task(const Task &T, int lb, float2 *tile, float2 *twl,
                                                 int *rml) {
  const int xbi = lb % T.cg.nxbf, t = lb / T.cg.nxbf;
  col_forward_fused<DELTA, P>(T.cg, xbi, t, T.ext, DELTA ? T.x : nullptr, T.b1, T.kcol, tile, twl,
                              rml);
}
template <class P>
__device__ __forceinline__ void col_adjoint_task(const Task &T, int lb, float2 *tile, float2 *twl,
                                                 int *rml) {
  const int xbi = lb % T.cg.nxbf, t = lb / T.cg.nxbf;
  col_adjoint_fused<P>(T.cg, xbi, t, T.zhat, T.b1, T.khz, tile, twl, rml);
}

// DELTA = false: K of the extragradient image (dual update);
// DELTA = true : K of ext - x (step rule, adapting tasks only).
// LDS entries of a column kind's twiddle and rowmap tables: compile-time
// plans are at most 256 long, narrow runtime plans 512, wide ones 1024.
__host__ __device__ constexpr int col_table_len(int kind) {
  return kind == 1 ? kColPlanMaxLen : (kind == 2 ? kMaxFftLen : kMaxFftLenNarrow);
}

template <bool DELTA, int KIND>
__global__ __launch_bounds__(NTC) void kt_col_forward(const Task *__restrict__ tasks, int n,
                                                      int blk0) {
  constexpr bool CT = KIND == 1;
  __shared__ float2 tile[kColTileElems];
  __shared__ float2 twl[col_table_len(KIND)];
  __shared__ int rml[col_table_len(KIND)];
  const int blk = blk0 + (int)blockIdx.x;
  const Task &T = tasks[find_task<kCol>(tasks, n, blk)];
  if (col_kind(T.cg) != KIND) return;
  if (!task_active(T.sc)) return;
  if (DELTA && !task_adapt(T.sc)) return;
  const int lb = blk - T.b_col;
  if constexpr (KIND == 0) {
    col_forward_task<DELTA, ColPlanRt>(T, lb, tile, twl, rml);
  } else if constexpr (KIND == 2) {
    col_forward_task<DELTA, ColPlanRtWide>(T, lb, tile, twl, rml);
  } else {
    switch (T.cg.cplan) {
      case kColPlan256: col_forward_task<DELTA, ColPlan256>(T, lb, tile, twl, rml); break;
      case kColPlan128: col_forward_task<DELTA, ColPlan128>(T, lb, tile, twl, rml); break;
      case kColPlan160: col_forward_task<DELTA, ColPlan160>(T, lb, tile, twl, rml); break;
      case kColPlan192: col_forward_task<DELTA, ColPlan192>(T, lb, tile, twl, rml); break;
      case kColPlan224: col_forward_task<DELTA, ColPlan224>(T, lb, tile, twl, rml); break;
      default: break;
    }
  }
}

// Row pass, one kernel per row kind (row_kind): 0 compile-time plans (16 or
// 32 rows per block), 1 narrow runtime plans, 2 wide runtime plans (RB rows
// per block).
template <int MODE, int RK>
__global__ __launch_bounds__(NT) void kt_row(const Task *__restrict__ tasks, int n, int blk0) {
  constexpr bool CT = RK == 0;
  __shared__ float2 tile[CT ? kFftTileElems : RTILE];
  __shared__ float2 tile2[(MODE == kRowDualUpdate && !CT) ? RTILE : 1];
  __shared__ float2 twl[CT ? kColPlanMaxLen : (RK == 2 ? kMaxFftLen : kMaxFftLenNarrow)];
  __shared__ double red[NT];
  __shared__ int mrl[CT ? kRowPlanMaxRows : kRowMaxRows];
  const int blk = blk0 + (int)blockIdx.x;
  const Task &T = tasks[find_task<kRow>(tasks, n, blk)];
  if (row_kind(T.cg) != RK) return;
  if (!task_active(T.sc)) return;
  if (MODE == kRowNorm2 && !task_adapt(T.sc)) return;
  const int lb = blk - T.b_row;
  double a = 0.0;
  if constexpr (CT) {
    if (T.cg.rplan == kRowPlan256)
      a = row_fused_ct<MODE, RowPlan256>(T.cg, lb, T.kcol, T.z, T.d, T.zhat, T.sc, tile, twl, mrl);
    else
      a = row_fused_ct<MODE, RowPlan128>(T.cg, lb, T.kcol, T.z, T.d, T.zhat, T.sc, tile, twl, mrl);
  } else {
    if (MODE == kRowDualUpdate)
      row_dual_fused<RK == 2>(T.cg, lb, T.kcol, T.z, T.d, T.zhat, T.sc, tile, tile2, twl, mrl);
    else
      a = row_norm2_fused<RK == 2>(T.cg, lb, T.kcol, tile, twl, mrl);
  }
  if (MODE == kRowNorm2) {
    const int tid = threadIdx.x;
    red[tid] = a;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    if (tid == 0) T.part_k[lb] = red[0];
  }
}

// The adjoint column pass keeps one accumulator per output element across
// the coil loop.  Runtime plans: asking for 4 waves per SIMD (the limit the
// LDS tile sets anyway) trims it from 130 to 128 VGPRs (col_adjoint 1.17 ->
// 0.98 ms, ICTGV2 256x256x30x16, 8 reconstructions).  Compile-time plans
// also hold the prefetched coil map values (3 waves per SIMD at 256; asking
// for 4 spills).
template <int KIND>
__device__ __forceinline__ void col_adjoint_block(const Task *__restrict__ tasks, int n, int blk0) {
  __shared__ float2 tile[kColTileElems];
  __shared__ float2 twl[col_table_len(KIND)];
  __shared__ int rml[col_table_len(KIND)];
  const int blk = blk0 + (int)blockIdx.x;
  const Task &T = tasks[find_task<kCol>(tasks, n, blk)];
  if (col_kind(T.cg) != KIND) return;
  if (!task_active(T.sc)) return;
  const int lb = blk - T.b_col;
  if constexpr (KIND == 0) {
    col_adjoint_task<ColPlanRt>(T, lb, tile, twl, rml);
  } else if constexpr (KIND == 2) {
    col_adjoint_task<ColPlanRtWide>(T, lb, tile, twl, rml);
  } else {
    switch (T.cg.cplan) {
      case kColPlan256: col_adjoint_task<ColPlan256>(T, lb, tile, twl, rml); break;
      case kColPlan128: col_adjoint_task<ColPlan128>(T, lb, tile, twl, rml); break;
      case kColPlan160: col_adjoint_task<ColPlan160>(T, lb, tile, twl, rml); break;
      case kColPlan192: col_adjoint_task<ColPlan192>(T, lb, tile, twl, rml); break;
      case kColPlan224: col_adjoint_task<ColPlan224>(T, lb, tile, twl, rml); break;
      default: break;
    }
  }
}
template <int KIND>
__global__ __launch_bounds__(NTC) __attribute__((amdgpu_waves_per_eu(4)))
void kt_col_adjoint_rt(const Task *__restrict__ tasks, int n, int blk0) {
  col_adjoint_block<KIND>(tasks, n, blk0);
}
template <int KIND>
__global__ __launch_bounds__(NTC) void kt_col_adjoint_ct(const Task *__restrict__ tasks, int n,
                                                         int blk0) {
  col_adjoint_block<KIND>(tasks, n, blk0);
}

// One block per task: step rule for adapting tasks, then the iteration
// counter.  The reductions run in a fixed order (deterministic).
__global__ __launch_bounds__(RT) void kt_tail(const Task *__restrict__ tasks, int n) {
  const Task &T = tasks[blockIdx.x];
  DevScalars *sc = T.sc;
  if (!task_active(sc)) return;
  if (task_adapt(sc)) {
    __shared__ double ra[RT], rb[RT];
    const int tid = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int j = tid; j < T.nvox_blocks; j += RT) {
      a += T.part_reg[2 * j];
      b += T.part_reg[2 * j + 1];
    }
    for (int j = tid; j < T.npart_k; j += RT) a += T.part_k[j];
    ra[tid] = a;
    rb[tid] = b;
    __syncthreads();
    for (int s = RT / 2; s > 0; s >>= 1) {
      if (tid < s) {
        ra[tid] += ra[tid + s];
        rb[tid] += rb[tid + s];
      }
      __syncthreads();
    }
    if (tid == 0) {
      apply_step_rule(sc, ra[0], rb[0]);
      if (T.sig_log) T.sig_log[sc->iter] = sc->sigma;
      if (T.tau_log) T.tau_log[sc->iter] = sc->tau;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) sc->iter += 1;
}

const char *kernel_name(int id) {
  static const char *names[kNumKernels] = {"dual", "col_forward", "row_dual_update", "col_adjoint",
                                           "primal", "step_reg", "col_forward_delta",
                                           "row_norm2", "tail"};
  return (id >= 0 && id < kNumKernels) ? names[id] : "?";
}

// Column-pass workgroups up to which the column pass is submitted before the
// dual: 16 per CU, i.e. up to four waves of the 4 resident per CU (the LDS
// tile).
static int col_resident_blocks() {
  static int n = [] {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
    return 16 * cus;
  }();
  return n;
}

bool default_col_first(const TaskGrid &gr) { return gr.col < col_resident_blocks(); }

// Column passes: one launch per kind of column plan present among the tasks.
template <bool DELTA, int K = 0>
static void launch_col_fwd(const TaskGrid &gr, const Task *d_tasks, int n, hipStream_t s) {
  if constexpr (K < kColKinds) {
    const int nb = gr.col_hi[K] - gr.col_lo[K];
    if (nb > 0)
      hipLaunchKernelGGL((kt_col_forward<DELTA, K>), dim3(nb), dim3(NTC), 0, s, d_tasks, n,
                         gr.col_lo[K]);
    launch_col_fwd<DELTA, K + 1>(gr, d_tasks, n, s);
  }
}
template <int K = 0>
static void launch_col_adj(const TaskGrid &gr, const Task *d_tasks, int n, hipStream_t s) {
  if constexpr (K < kColKinds) {
    const int nb = gr.col_hi[K] - gr.col_lo[K];
    if (nb > 0) {
      if constexpr (K != 1)
        hipLaunchKernelGGL((kt_col_adjoint_rt<K>), dim3(nb), dim3(NTC), 0, s, d_tasks, n,
                           gr.col_lo[K]);
      else
        hipLaunchKernelGGL((kt_col_adjoint_ct<K>), dim3(nb), dim3(NTC), 0, s, d_tasks, n,
                           gr.col_lo[K]);
    }
    launch_col_adj<K + 1>(gr, d_tasks, n, s);
  }
}

template <int MODE, int RK = 0>
static void launch_row(const TaskGrid &gr, const Task *d_tasks, int n, hipStream_t s) {
  if constexpr (RK < kRowKinds) {
    const int nb = gr.row_hi[RK] - gr.row_lo[RK];
    if (nb > 0)
      hipLaunchKernelGGL((kt_row<MODE, RK>), dim3(nb), dim3(NT), 0, s, d_tasks, n, gr.row_lo[RK]);
    launch_row<MODE, RK + 1>(gr, d_tasks, n, s);
  }
}

// Dual (primal = false) or primal update over every task, in the grid's
// launch variant (TaskGrid::lanes).
template <int M, typename DT, int F>
static void launch_dp(bool primal, int nblk, hipStream_t s, const Task *t, int n) {
  if (primal)
    hipLaunchKernelGGL((kt_primal<M, DT, F>), dim3(nblk), dim3(RT), 0, s, t, n);
  else
    hipLaunchKernelGGL((kt_dual<M, DT, F>), dim3(nblk), dim3(RT), 0, s, t, n);
}
template <typename DT, int F>
static void launch_dp_m(int method, bool primal, int nblk, hipStream_t s, const Task *t, int n) {
  if (method == kTV) launch_dp<kTV, DT, F>(primal, nblk, s, t, n);
  else if (method == kTGV2) launch_dp<kTGV2, DT, F>(primal, nblk, s, t, n);
  else if (method == kICTGV2) launch_dp<kICTGV2, DT, F>(primal, nblk, s, t, n);
  else launch_dp<kICTV, DT, F>(primal, nblk, s, t, n);
}
template <int F>
static void launch_dp_v(int method, int ds, bool primal, int nblk, hipStream_t s, const Task *t,
                        int n) {
  if (ds == kDualF32) launch_dp_m<float2, F>(method, primal, nblk, s, t, n);
  else if (ds == kDualBF16) launch_dp_m<bf16x2, F>(method, primal, nblk, s, t, n);
  else launch_dp_m<q16x2, F>(method, primal, nblk, s, t, n);
}
static void launch_dual_primal(bool primal, int method, int ds, const TaskGrid &gr, hipStream_t s,
                               const Task *t, int n) {
  if (gr.lanes == 1) {
    if (gr.vox > 0) launch_dp_v<kVox>(method, ds, primal, gr.vox, s, t, n);
    return;
  }
  if (gr.dp2 > 0) launch_dp_v<kDp2>(method, ds, primal, gr.dp2, s, t, n);
  if (gr.dp1 > 0) launch_dp_v<kDp1>(method, ds, primal, gr.dp1, s, t, n);
}

void launch_kernel(int id, int method, int ds, const Task *d_tasks, int n, const TaskGrid &gr,
                   hipStream_t s) {
  if (n <= 0) return;
  const dim3 bRT(RT), bNT(NT), bNC(NTC);
  switch (id) {
    case kKDual: launch_dual_primal(false, method, ds, gr, s, d_tasks, n); break;
    case kKColFwd: launch_col_fwd<false>(gr, d_tasks, n, s); break;
    case kKRowDual: launch_row<kRowDualUpdate>(gr, d_tasks, n, s); break;
    case kKColAdj: launch_col_adj(gr, d_tasks, n, s); break;
    case kKPrimal: launch_dual_primal(true, method, ds, gr, s, d_tasks, n); break;
    case kKStepReg:
      if (gr.vox) {
        if (method == kTV) hipLaunchKernelGGL(kt_step_reg<kTV>, dim3(gr.vox), bRT, 0, s, d_tasks, n);
        else if (method == kTGV2) hipLaunchKernelGGL(kt_step_reg<kTGV2>, dim3(gr.vox), bRT, 0, s, d_tasks, n);
        else if (method == kICTGV2) hipLaunchKernelGGL(kt_step_reg<kICTGV2>, dim3(gr.vox), bRT, 0, s, d_tasks, n);
        else hipLaunchKernelGGL(kt_step_reg<kICTV>, dim3(gr.vox), bRT, 0, s, d_tasks, n);
      }
      break;
    case kKColFwdDelta: launch_col_fwd<true>(gr, d_tasks, n, s); break;
    case kKRowNorm2: launch_row<kRowNorm2>(gr, d_tasks, n, s); break;
    case kKTail: hipLaunchKernelGGL(kt_tail, dim3(n), bRT, 0, s, d_tasks, n); break;
    default: break;
  }
  AV_LAUNCH_CHECK();
}

void launch_iteration(int method, int ds, const Task *d_tasks, int n, const TaskGrid &gr,
                      bool with_step, const IterStreams &st, const IterHooks &hooks) {
  if (n <= 0) return;
  hipStream_t s = st.main;
  const bool two = st.side != nullptr;
  hipStream_t s2 = two ? st.side : st.main;
  auto fork = [&]() {
    if (!two) return;
    AV_HIP(hipEventRecord(st.fork, s));
    AV_HIP(hipStreamWaitEvent(s2, st.fork, 0));
  };
  auto join = [&]() {
    if (!two) return;
    AV_HIP(hipEventRecord(st.join, s2));
    AV_HIP(hipStreamWaitEvent(s, st.join, 0));
  };
  const dim3 bRT(RT), bNT(NT), bNC(NTC);
  // dual update (reads ext, writes y) beside the data-term chain (reads ext,
  // writes kcol / z / zhat / khz); the primal update needs both
  const bool col_first = gr.col_first >= 0 ? gr.col_first == 1 : default_col_first(gr);
  if (two && col_first) {
    // The column pass of up to a few waves of workgroups (one reconstruction
    // of 256x256x30x16: 480 workgroups; eight: 3840) is submitted first, so
    // that its workgroups start and the dual fills the rest of the CUs; the
    // dual still depends only on what precedes the column pass.  Larger
    // column passes go after the dual (measured on MI355X, ICTGV2
    // 256x256x30x16: 1 reconstruction 2507 -> 2724 it/s with this order;
    // 8 reconstructions 3268 -> 3283 it/s once the dual reads and writes y
    // with non-temporal accesses, 3219 -> 3172 before).
    AV_HIP(hipEventRecord(st.fork, s));
    launch_col_fwd<false>(gr, d_tasks, n, s);
    AV_HIP(hipStreamWaitEvent(s2, st.fork, 0));
    launch_dual_primal(false, method, ds, gr, s2, d_tasks, n);
  } else {
    fork();
    launch_dual_primal(false, method, ds, gr, s2, d_tasks, n);
    launch_col_fwd<false>(gr, d_tasks, n, s);
  }
  AV_LAUNCH_CHECK();
  launch_row<kRowDualUpdate>(gr, d_tasks, n, s);
  launch_col_adj(gr, d_tasks, n, s);
  AV_LAUNCH_CHECK();
  if (hooks.fn) hooks.fn(hooks.ctx, 0, s);
  join();
  launch_dual_primal(true, method, ds, gr, s, d_tasks, n);
  AV_LAUNCH_CHECK();
  if (with_step) {
    // regulariser sums (read ext, x) beside K(ext - x); the tail needs both
    fork();
    if (gr.vox > 0) {
      if (method == kTV) hipLaunchKernelGGL(kt_step_reg<kTV>, dim3(gr.vox), bRT, 0, s2, d_tasks, n);
      else if (method == kTGV2) hipLaunchKernelGGL(kt_step_reg<kTGV2>, dim3(gr.vox), bRT, 0, s2, d_tasks, n);
      else if (method == kICTGV2) hipLaunchKernelGGL(kt_step_reg<kICTGV2>, dim3(gr.vox), bRT, 0, s2, d_tasks, n);
      else hipLaunchKernelGGL(kt_step_reg<kICTV>, dim3(gr.vox), bRT, 0, s2, d_tasks, n);
    }
    AV_LAUNCH_CHECK();
    launch_col_fwd<true>(gr, d_tasks, n, s);
    launch_row<kRowNorm2>(gr, d_tasks, n, s);
    AV_LAUNCH_CHECK();
    if (hooks.fn) hooks.fn(hooks.ctx, 1, s);
    join();
  }
  hipLaunchKernelGGL(kt_tail, dim3(n), bRT, 0, s, d_tasks, n);
  AV_LAUNCH_CHECK();
}

}  // namespace avionic
