// Task-list execution of the PD iteration.
//
// One Task is one reconstruction's device state.  Every hot kernel takes an
// array of tasks and a flattened grid: block b belongs to the task whose first
// block (per kernel family) is the largest not exceeding b.  A launch therefore
// advances any number of reconstructions of one method at once, each with its
// own shape, step sizes, iteration counter and schedule; a single
// reconstruction is the one-task case of the same kernels.
//
// Per-task predicates evaluated on the device:
//   active: sc->iter < sc->maxit        (every kernel; a finished task idles)
//   adapt : schedule == 1 || (schedule == 0 && (iter < 10 || iter % 50 == 0));
//           schedule 2 keeps fixed steps
//           (step-size kernels; tv.cpp:172, tgv2.cpp:306, ictgv2.cpp:448)
// so the same captured graph can be replayed for any mix of iteration counts.
#pragma once

#include "cartesian.hpp"
#include "common.hpp"
#include "reg.hpp"

namespace avionic {

struct Task {
  Grid3 g;
  RegParams rp;
  CartGeom cg;
  float2 *x, *ext;     // primal, P components each
  void *y;             // dual, D components (float2, bf16x2 or q16x2)
  float2 *khz;         // K^H z (image)
  float2 *z;           // compact k-space dual
  const float2 *d;     // compact data
  float2 *kcol, *zhat; // operator scratch (compact)
  const float2 *b1;
  DevScalars *sc;
  double *sig_log, *tau_log;  // indexed by iteration
  double *part_reg;    // 2 per voxel block
  double *part_k;      // K-norm partials of the step rule (npart_k of them)
  int nvox_blocks, ncol_blocks, nrow_blocks;
  // Dual / primal update blocks when a launch runs pairs of voxels
  // (TaskGrid::lanes 2): a task of even width runs ndp2_blocks = pair_blocks
  // blocks of a pair of voxels of a row per lane (ndp1_blocks = 0), a task of
  // odd width its ndp1_blocks = nvox_blocks one-voxel blocks (ndp2_blocks =
  // 0).  With lanes 1 every task runs its nvox_blocks one-voxel blocks.
  int ndp1_blocks, ndp2_blocks;
  int npart_k;                // Cartesian: nrow_blocks; other operators: their own
  int b_vox, b_col, b_row;    // first block of this task per kernel family
  int b_dp1, b_dp2;
  int schedule;               // 0 reference schedule, 1 every iteration
};

// Launch one iteration's kernels for `n` tasks of one method / dual storage.
// tasks: device array; the host mirror gives the grid sizes.
struct TaskGrid {
  int vox = 0, col = 0, row = 0;
  // Dual / primal blocks of the one-voxel-per-lane and the pair tasks.
  int dp1 = 0, dp2 = 0;
  // Launch variant (both give the same bits): lanes 2 runs pairs of voxels
  // for the tasks of even width, lanes 1 one voxel per lane for every task;
  // col_first 1 submits the column pass before the dual update, 0 after it,
  // -1 decides by the column pass's grid size.
  int lanes = 2, col_first = -1;
  // Blocks [lo, hi) spanning the tasks of each column kind (col_kind) and of
  // each row kind (row_kind); the engine orders its tasks by kind, so each
  // span holds only that kind's blocks.
  int col_lo[kColKinds] = {}, col_hi[kColKinds] = {};
  int row_lo[kRowKinds] = {}, row_hi[kRowKinds] = {};
};

// Streams of one iteration: the regulariser kernels that do not depend on
// the data-term chain (the dual update; the step rule's regulariser sums)
// run on `side`, forked from and joined back into `main` with the two
// events, so that the HBM-bound stencil kernel overlaps the latency-bound
// FFT passes.  side == nullptr runs everything on main.
struct IterStreams {
  hipStream_t main = nullptr, side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
// Work of tasks whose operator has no task-list passes (non-Cartesian
// operators, ncol_blocks = nrow_blocks = 0): fn(ctx, 0, main) runs their
// data-term dual update after the Cartesian z chain, fn(ctx, 1, main) their
// step-rule K norms after the Cartesian ones; both before the join.
struct IterHooks {
  void (*fn)(void *ctx, int phase, hipStream_t s) = nullptr;
  void *ctx = nullptr;
};
void launch_iteration(int method, int dual_storage, const Task *d_tasks, int n, const TaskGrid &grid,
                      bool with_step, const IterStreams &st, const IterHooks &hooks = IterHooks{});
// The column-pass order col_first = -1 resolves to: the column pass first
// when its grid is at most a few waves of workgroups on this device.
bool default_col_first(const TaskGrid &grid);

// Kernel families of one iteration, in launch order.
enum KernelId {
  kKDual = 0, kKColFwd, kKRowDual, kKColAdj, kKPrimal, kKStepReg, kKColFwdDelta, kKRowNorm2,
  kKTail, kNumKernels
};
const char *kernel_name(int id);
// Launch one kernel family alone (for timing).
void launch_kernel(int id, int method, int dual_storage, const Task *d_tasks, int n, const TaskGrid &grid,
                   hipStream_t s);

}  // namespace avionic
